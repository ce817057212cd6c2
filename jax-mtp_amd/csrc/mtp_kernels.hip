// HIP kernels for one MTP force evaluation on gfx950 (CDNA4).
//
// Replaces the XLA program exported from jax_backend.py:374-460
// (calc_energy_forces_stress_padded): per-atom energy and its gradient with
// respect to every neighbor vector (jax_backend.py:20-106, value_and_grad),
// the Newton-3 force scatter (:263-303), the stress (:355-370) and the masked
// energy sum (:431-456).  Everything computes in FP64 except, optionally,
// the radial basis (Chebyshev terms and f_mu), which the radial-FP32 variant
// evaluates in FP32.
//
// Kernel map:
//   atom_kernel     one 256-thread workgroup per atom row.  LDS holds the
//                   atom's moments m[amc] and adjoints dm[amc] in FP64.
//                     A  radial basis per slot, compaction of valid slots
//                     B  basic moments  m[mu,alpha] = sum_k f_mu(r_k) u_k^alpha
//                     C  contraction layers (alpha_index_times, by depth)
//                     D  E_i = species_coeff + xi . m[mapping]
//                     E  reverse layers: dm = dE/dm (gather form)
//                     F  g_k = dE_i/dr_k from dm of the basic moments
//                     G  per-atom virial W_i = sum_k r_k (x) g_k
//   zero/count/scan/fill/force kernels: deterministic Newton-3 force gather
//   totals_partial_kernel + totals_kernel: E = sum_{i<n_e} E_i, W = sum_i W_i
//                   -> stress, virial (two-stage, deterministic)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mtp_device.hpp"
#include "mtp_kernels.hpp"

namespace mtp {

constexpr int kBlock = kAtomBlock;

// A slot vector with a NaN or Inf component turns the reference's energy and
// forces into NaN (r, u = r_ij / r and the Chebyshev terms of such a slot are
// not finite, and s = 0 does not mask them, jax_backend.py:83-95, 232).  The
// kernels drop such a slot like an invalid one and raise the status word
// instead, which the totals turn into a NaN energy and status bit 2 (error
// MTP_ERR_NONFINITE at the C ABI).  Called by every lane of the wave.
__device__ __forceinline__ void flag_nonfinite(int* status, bool bad)
{
    const uint64_t m = __ballot(bad);
    if (status && m != 0ull && (threadIdx.x & (kWave - 1)) == 0) atomicOr(status, 1);
}

// LDS layout (in doubles) shared by host sizing and the kernel.  gm: the
// moment and adjoint regions (mlen + dlen doubles) live in a global slice.
__host__ __device__ inline AtomLds atom_lds_layout(int amc, int nslots, int R, int nu_max, int n_parts,
                                                   bool gm = false)
{
    AtomLds L;
    int mlen = amc;
    if (mlen < n_parts * nslots * 3) mlen = n_parts * nslots * 3;
    int dlen = amc;
    if (dlen < nslots * 3 * (nu_max + 1)) dlen = nslots * 3 * (nu_max + 1);
    L.mlen = mlen;
    L.dlen = dlen;
    L.m = 0;
    L.dm = gm ? 0 : L.m + mlen;
    L.f = gm ? 0 : L.dm + dlen;
    L.fp = L.f + R * nslots;
    L.vx = L.fp + R * nslots;
    L.vy = L.vx + nslots;
    L.vz = L.vy + nslots;
    L.r = L.vz + nslots;
    L.red = L.r + nslots;
    L.ints = L.red + kBlock;                 // int region starts here (in doubles)
    L.total_doubles = L.ints + (2 * nslots + 8 + 1) / 2 + 1;
    return L;
}

size_t atom_lds_bytes(int amc, int nslots, int R, int nu_max, int n_parts, bool global_moments)
{
    return sizeof(double) * (size_t)atom_lds_layout(amc, nslots, R, nu_max, n_parts, global_moments).total_doubles;
}

size_t atom_global_moment_doubles(int amc, int nslots, int nu_max, int n_parts)
{
    const AtomLds L = atom_lds_layout(amc, nslots, 1, nu_max, n_parts, true);
    return (size_t)L.mlen + (size_t)L.dlen;
}

// One row of the per-target kernel.  gm: the moments and adjoints in the
// global slice gm[0, mlen + dlen) of this workgroup, else in LDS.
template <typename RadialReal>
__device__ __forceinline__ void atom_row_v1(const DevPlan& P, const AtomArgs& A, const int row, double* lds,
                                            const AtomLds& L, double* gm)
{
    const int tid = threadIdx.x;
    const int K = A.nslots;
    double* m = gm ? gm : lds + L.m;
    double* dm = gm ? gm + L.mlen : lds + L.dm;
    double* f = lds + L.f;
    double* fp = lds + L.fp;
    double* vx = lds + L.vx;
    double* vy = lds + L.vy;
    double* vz = lds + L.vz;
    double* rr = lds + L.r;
    double* red = lds + L.red;
    int* ints = reinterpret_cast<int*>(lds + L.ints);
    int* slot_of = ints;          // [K] compact index -> slot
    int* cidx = ints + K;         // [K] slot -> compact index or -1
    int* wave_cnt = ints + 2 * K; // [4]
    int* nvp = wave_cnt + 4;

    const int ti = A.itypes[row];
    const double* rij_row = A.rij + (size_t)row * K * 3;
    const int* jt_row = A.jtypes + (size_t)row * K;

    // ---- A: radial basis and compaction of valid slots (slot order kept).
    // The host guarantees K <= kBlock.
    {
        const int k = tid;
        double v0 = 0, v1 = 0, v2 = 0, r = 1.0;
        int tk = -1;
        bool valid = false;
        if (k < K) {
            v0 = rij_row[3 * k]; v1 = rij_row[3 * k + 1]; v2 = rij_row[3 * k + 2];
            tk = jt_row[k];
            r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
            valid = (tk >= 0) && (r < P.max_dist);
        }
        flag_nonfinite(A.status, k < K && !isfinite(r));
        const uint64_t mask = __ballot(valid);
        const int lane = tid & (kWave - 1), wave = tid / kWave;
        if (lane == 0) wave_cnt[wave] = __popcll(mask);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
        if (k < K) cidx[k] = valid ? pos : -1;
        if (valid) {
            double fv[kMaxR], dfv[kMaxR];
            radial_slot<RadialReal>(P, r, ti, tk, fv, dfv);
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu)
                if (mu < P.R) { f[mu * K + pos] = fv[mu]; fp[mu * K + pos] = dfv[mu]; }
            vx[pos] = v0; vy[pos] = v1; vz[pos] = v2; rr[pos] = r;
            slot_of[pos] = k;
        }
        if (tid == 0) {
            int tot = 0;
            for (int w = 0; w < kBlock / kWave; ++w) tot += wave_cnt[w];
            *nvp = tot;
        }
        __syncthreads();
    }
    const int nv = *nvp;
    for (int x = P.B + tid; x < P.amc; x += kBlock) m[x] = 0.0;

    // ---- B: basic moments.  Power table u^p in the dm region.
    const int np1 = P.nu_max + 1;
    double* pw = dm;
    for (int t = tid; t < nv * 3; t += kBlock) {
        const int k = t / 3, q = t - 3 * k;
        const double inv = 1.0 / rr[k];
        const double u = (q == 0 ? vx[k] : (q == 1 ? vy[k] : vz[k])) * inv;
        double* row_pw = pw + (size_t)t * np1;
        double v = 1.0;
        row_pw[0] = 1.0;
        for (int p = 1; p < np1; ++p) { v *= u; row_pw[p] = v; }
    }
    __syncthreads();
    for (int t = tid; t < P.n_mono; t += kBlock) {
        const I4 mo = P.mono[t];
        const int nu = mo.w & 0xff, alpha = (mo.w >> 8) & 0xffff;
        const int mumax = P.mu_max_of_nu[nu];
        double acc[kMaxR];
#pragma unroll
        for (int mu = 0; mu < kMaxR; ++mu) acc[mu] = 0.0;
        for (int k = 0; k < nv; ++k) {
            const double* pk = pw + (size_t)k * 3 * np1;
            const double Pm = pk[mo.x] * pk[np1 + mo.y] * pk[2 * np1 + mo.z];
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu)
                if (mu <= mumax) acc[mu] += f[mu * K + k] * Pm;
        }
#pragma unroll
        for (int mu = 0; mu < kMaxR; ++mu)
            if (mu <= mumax) m[P.blk_off[mu][nu] + alpha] = acc[mu];
    }
    __syncthreads();

    // ---- C: contraction layers
    for (int d = 1; d <= P.n_depth; ++d) {
        const int t0 = P.fwd_layer[d], t1 = P.fwd_layer[d + 1];
        for (int t = t0 + tid; t < t1; t += kBlock) {
            double s = 0.0;
            const int e1 = P.fwd_beg[t + 1];
            for (int e = P.fwd_beg[t]; e < e1; ++e) {
                const uint32_t pr = P.fwd_pair[e];
                s += P.fwd_mult[e] * m[pr & 0xffffu] * m[pr >> 16];
            }
            m[P.fwd_tgt[t]] = s;
        }
        __syncthreads();
    }

    // ---- D: site energy
    {
        double acc = 0.0;
        for (int s = tid; s < P.S; s += kBlock) acc += P.moment_coeffs[s] * m[P.mapping[s]];
        red[tid] = acc;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        if (tid == 0) A.Ei[row] = P.species_coeffs[wrap_species(ti, P.Sp)] + red[0];
    }

    // ---- E: reverse layers (deepest first), gather form
    for (int li = 0; li <= P.n_depth; ++li) {
        const int t0 = P.rev_layer[li], t1 = P.rev_layer[li + 1];
        for (int t = t0 + tid; t < t1; t += kBlock) {
            const int x = P.rev_x[t];
            double s = P.xi_of[x];
            const int e1 = P.rev_beg[t + 1];
            for (int e = P.rev_beg[t]; e < e1; ++e) {
                const uint32_t pr = P.rev_pair[e];
                s += P.rev_mult[e] * m[pr & 0xffffu] * dm[pr >> 16];
            }
            dm[x] = s;
        }
        __syncthreads();
    }

    // ---- F: per-slot gradient g_k = dE_i/dr_k.  Thread = (slot, part).
    double* gpart = m;  // m is no longer needed
    {
        const int kk = tid & (kWave - 1), part = tid / kWave;
        for (int k = kk; k < nv && part < P.n_parts; k += kWave) {
            const double ir = 1.0 / rr[k];
            const double ux = vx[k] * ir, uy = vy[k] * ir, uz = vz[k] * ir;
            double px[kMaxNu + 1], py[kMaxNu + 1], pz[kMaxNu + 1];
            px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
            for (int p = 1; p <= kMaxNu; ++p) { px[p] = px[p - 1] * ux; py[p] = py[p - 1] * uy; pz[p] = pz[p - 1] * uz; }
            double sdf = 0.0, sfn = 0.0, Gx = 0.0, Gy = 0.0, Gz = 0.0;
            const int b1 = P.part_beg[part + 1];
            for (int bi = P.part_beg[part]; bi < b1; ++bi) {
                const I2 bl = P.part_blk[bi];
                const int mu = bl.x & 0xff, nu = bl.y;
                const double* d = dm + (bl.x >> 8);
                const double fm = f[mu * K + k], fpm = fp[mu * K + k];
                if (nu == 0) {
                    sdf += fpm * d[0];
                } else {
                    double Hx = 0.0, Hy = 0.0, Hz = 0.0;
                    grad_block_dispatch(nu, d, px, py, pz, Hx, Hy, Hz);
                    const double uH = ux * Hx + uy * Hy + uz * Hz;
                    sdf += fpm * uH * recip_order(nu);
                    sfn += fm * uH;
                    Gx += fm * Hx; Gy += fm * Hy; Gz += fm * Hz;
                }
            }
            double* gp = gpart + ((size_t)part * K + k) * 3;
            gp[0] = sdf * ux + (Gx - sfn * ux) * ir;
            gp[1] = sdf * uy + (Gy - sfn * uy) * ir;
            gp[2] = sdf * uz + (Gz - sfn * uz) * ir;
        }
    }
    __syncthreads();
    double* gk = dm;  // dm no longer needed
    for (int t = tid; t < nv * 3; t += kBlock) {
        double s = 0.0;
        for (int part = 0; part < P.n_parts; ++part) s += gpart[(size_t)part * K * 3 + t];
        gk[t] = s;
    }
    __syncthreads();
    double* g_row = A.g + (size_t)row * K * 3;
    for (int t = tid; t < K * 3; t += kBlock) {
        const int k = t / 3, q = t - 3 * k;
        const int c = cidx[k];
        g_row[t] = c >= 0 ? gk[3 * c + q] : 0.0;
    }
    // ---- G: per-atom virial W_i[a][b] = sum_k r_k[a] g_k[b]
    if (tid < 9) {
        const int a = tid / 3, b = tid - 3 * (tid / 3);
        const double* va = a == 0 ? vx : (a == 1 ? vy : vz);
        double s = 0.0;
        for (int k = 0; k < nv; ++k) s += va[k] * gk[3 * k + b];
        A.Wi[(size_t)row * 9 + tid] = s;
    }
}

// GM false: one workgroup per row (grid = rows).  GM true: a grid of
// P.mscratch_grid workgroups walks the rows, each with its global slice
// (barrier between rows: the next row reuses the LDS and the slice).
template <typename RadialReal, bool GM>
__global__ void __launch_bounds__(kBlock) atom_kernel(const DevPlan P, AtomArgs A)
{
    extern __shared__ double lds[];
    const AtomLds L = atom_lds_layout(P.amc, A.nslots, P.R, P.nu_max, P.n_parts, GM);
    if (!GM) {
        atom_row_v1<RadialReal>(P, A, blockIdx.x, lds, L, nullptr);
        return;
    }
    double* gm = P.mscratch + (size_t)blockIdx.x * (size_t)P.mscratch_stride;
    for (int row = blockIdx.x; row < A.nrows; row += gridDim.x) {
        atom_row_v1<RadialReal>(P, A, row, lds, L, gm);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// atom_kernel_v2: NT threads per atom and balanced contraction layers.
//
// Each layer is a list of host-planned work items of about (entries / NT)
// entries, one per thread: runs of whole target moments, or chunks of one
// long target whose partial sums are combined afterwards in a fixed order
// (deterministic, no atomics).  Item tables are transposed so that the
// threads of a wave read consecutive 8-byte entries (coalesced), and each
// entry packs both operand indices and the multiplier.
// ---------------------------------------------------------------------------
size_t atom_lds2_bytes(int amc, int n_live, int B, int nslots, int R, int nu_max, int n_parts, int nt,
                       int max_partials, int xi_len, bool tables, int coef_len)
{
    return sizeof(double) * (size_t)atom_lds2_layout(amc, n_live, B, nslots, R, nu_max, n_parts, nt, max_partials,
                                                     xi_len, tables, coef_len)
                                .total_doubles;
}

hipError_t set_atom_kernel_lds_exact(int variant, size_t bytes);

// Operand addresses of an entry word x = a | b << 16: byte offsets from
// the bases ab and bb, or (ABS) 8-byte slot indices from the LDS base lb.
template <bool ABS>
__device__ __forceinline__ const double* stream_op_a(uint32_t x, const char* ab, const char* lb)
{
    return reinterpret_cast<const double*>(ABS ? lb + ((x & 0xffffu) << 3) : ab + (x & 0xffffu));
}
template <bool ABS>
__device__ __forceinline__ const double* stream_op_b(uint32_t x, const char* bb, const char* lb)
{
    return reinterpret_cast<const double*>(ABS ? lb + ((x >> 16) << 3) : bb + (x >> 16));
}

// Store v at LDS byte offset off.  The atom kernels have no static LDS, so
// the dynamic array starts at LDS address 0 and off is the address itself
// (no add of the array's base per store).
__device__ __forceinline__ void lds_store(uint32_t off, double v)
{
    *reinterpret_cast<__attribute__((address_space(3))) double*>(off) = v;
}

template <bool ABS>
__device__ __forceinline__ void stream_steps(const uint4* q, const char* ab, const char* bb, char* lb, double& acc)
{
    constexpr int U = kStreamUnroll;
    double a0[U], b0[U], a1[U], b1[U], a2[U], b2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        a0[u] = *stream_op_a<ABS>(q[u].x, ab, lb);
        b0[u] = *stream_op_b<ABS>(q[u].x, bb, lb);
        a1[u] = *stream_op_a<ABS>(q[u].y, ab, lb);
        b1[u] = *stream_op_b<ABS>(q[u].y, bb, lb);
        a2[u] = *stream_op_a<ABS>(q[u].z, ab, lb);
        b2[u] = *stream_op_b<ABS>(q[u].z, bb, lb);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        // three entries with one integer multiplier; on the last step of
        // the lane's item (per lane) the sum goes to the item's destination
        // and restarts from zero
        double t = a0[u] * b0[u];
        t = fma(a1[u], b1[u], t);
        t = fma(a2[u], b2[u], t);
        const uint32_t w = q[u].w;
        acc = fma((double)stream_mult(w), t, acc);
        const uint32_t dst = stream_dest(w);
        if (dst) {
            lds_store(dst, acc);
            acc = 0.0;
        }
    }
}

// A wave's entry stream of one layer: nsteps (a multiple of kStreamUnroll)
// steps from tb (this lane's column), table loads one batch ahead.
template <bool ABS>
__device__ __forceinline__ void stream_walk(const uint4* tb, int nsteps, const char* bb, char* lb)
{
    constexpr int U = kStreamUnroll;
    double acc = 0.0;
    uint4 qa[U], qb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) qa[u] = tb[(size_t)u * kWave];
    for (int s0 = 0;;) {
        if (s0 + U < nsteps) {
#pragma unroll
            for (int u = 0; u < U; ++u) qb[u] = tb[(size_t)(s0 + U + u) * kWave];
        }
        stream_steps<ABS>(qa, lb, bb, lb, acc);
        s0 += U;
        if (s0 >= nsteps) break;
        if (s0 + U < nsteps) {
#pragma unroll
            for (int u = 0; u < U; ++u) qa[u] = tb[(size_t)(s0 + U + u) * kWave];
        }
        stream_steps<ABS>(qb, lb, bb, lb, acc);
        s0 += U;
        if (s0 >= nsteps) break;
    }
}

template <int NT, bool REV>
__device__ __forceinline__ void contraction_layer(const DevPlan& P, int L, double* lds, double* dm, double* part,
                                                  int tid, const double* xi)
{
    constexpr int NW = NT / kWave;
    const int lane = tid & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
    const I4 wv = (REV ? P.rev_wave : P.fwd_wave)[L * NW + w];
    const int nsteps = wv.y;
    // This thread's first row of the layer's combination and (reverse) its
    // first base-only and xi-added rows: their table entries and xi values
    // are global loads, issued here so that the stream walk hides their
    // latency (after the walk's barrier they would be two dependent global
    // round trips on every layer's critical path).
    const int c0 = REV ? P.rev_cb_beg[L] : P.fwd_cb_beg[L];
    const int c1 = REV ? P.rev_cb_beg[L + 1] : P.fwd_cb_beg[L + 1];
    const I4* comb = REV ? P.rev_comb : P.fwd_comb;
    const int cA = c0 + tid;
    I4 cbA = {0, 0, 0, 0};
    double xA = 0.0;
    if (cA < c1) {
        cbA = comb[cA];
        if (REV) xA = xi[cbA.x];
    }
    int boX = -1, xaX = -1;
    double boV = 0.0, xaV = 0.0;
    if (REV) {
        const int tb = P.rev_bo_beg[L] + tid, ta = P.rev_xa_beg[L] + tid;
        if (tb < P.rev_bo_beg[L + 1]) { boX = P.rev_bo[tb]; boV = xi[boX]; }
        if (ta < P.rev_xa_beg[L + 1]) { xaX = P.rev_xa[ta]; xaV = xi[xaX]; }
    }
    if (nsteps > 0) {
        const uint4* tb = reinterpret_cast<const uint4*>(REV ? P.rev_tbl : P.fwd_tbl) + (size_t)wv.x * kWave + lane;
        char* lb = reinterpret_cast<char*>(lds);
        const char* bb = REV ? reinterpret_cast<const char*>(dm) : lb;
        if (P.abs_slots) stream_walk<true>(tb, nsteps, bb, lb);
        else stream_walk<false>(tb, nsteps, bb, lb);
    }
    __syncthreads();
    double* dst = REV ? dm : lds + kLdsM;
    if (cA < c1) {
        double acc = xA;
        for (int p = cbA.y; p < cbA.z; ++p) acc += part[p];
        dst[cbA.x] = acc;
    }
    for (int c = cA + NT; c < c1; c += NT) {
        const I4 cb = comb[c];
        double acc = REV ? xi[cb.x] : 0.0;
        for (int p = cb.y; p < cb.z; ++p) acc += part[p];
        dst[cb.x] = acc;
    }
    if (REV) {
        if (boX >= 0) dm[boX] = boV;
        for (int t = P.rev_bo_beg[L] + tid + NT; t < P.rev_bo_beg[L + 1]; t += NT) {
            const int x = P.rev_bo[t];
            dm[x] = xi[x];
        }
        if (xaX >= 0) dm[xaX] += xaV;
        for (int t = P.rev_xa_beg[L] + tid + NT; t < P.rev_xa_beg[L + 1]; t += NT) {
            const int x = P.rev_xa[t];
            dm[x] += xi[x];
        }
    }
    if (c1 > c0 || REV) __syncthreads();
}

// Gradient of one slot k over the (mu, nu) blocks of part prt: g_k from the
// adjoints dm of the basic moments.  H = sum_alpha dm_alpha grad_u u^alpha
// over the monomials of degree nu - 1 (power tables of PN orders, enough
// for nu <= PN), u . H = nu sum_alpha dm_alpha u^alpha (Euler).  The
// adjoints of this row are dm[DS * x].
template <int PN, int DS = 1>
__device__ __forceinline__ void slot_gradient(const DevPlan& P, int prt, int k, int K, const double* dm,
                                              const double* gf, const double* gfp, const double* vx,
                                              const double* vy, const double* vz, const double* rr, double* gpart)
{
    const double ir = 1.0 / rr[k];
    const double ux = vx[k] * ir, uy = vy[k] * ir, uz = vz[k] * ir;
    double px[PN], py[PN];
    px[0] = py[0] = 1.0;
#pragma unroll
    for (int p = 1; p < PN; ++p) { px[p] = px[p - 1] * ux; py[p] = py[p - 1] * uy; }
    double sdf = 0.0, sfn = 0.0, Gx = 0.0, Gy = 0.0, Gz = 0.0;
    const int b1 = P.part_beg[prt + 1];
    for (int bi = P.part_beg[prt]; bi < b1; ++bi) {
        const I2 bl = P.part_blk[bi];
        const int mu = bl.x & 0xff, nu = bl.y;
        const double* d = dm + DS * (bl.x >> 8);
        const double fm = gf[mu * K + k], fpm = gfp[mu * K + k];
        if (nu == 0) {
            sdf += fpm * d[0];
        } else {
            double Hx = 0.0, Hy = 0.0, Hz = 0.0;
            grad_block_dispatch_xy<PN, DS>(nu, d, px, py, uz, Hx, Hy, Hz);
            const double uH = ux * Hx + uy * Hy + uz * Hz;
            sdf += fpm * uH * recip_order(nu);
            sfn += fm * uH;
            Gx += fm * Hx; Gy += fm * Hy; Gz += fm * Hz;
        }
    }
    double* gp = gpart + ((size_t)prt * K + k) * 3;
    gp[0] = sdf * ux + (Gx - sfn * ux) * ir;
    gp[1] = sdf * uy + (Gy - sfn * uy) * ir;
    gp[2] = sdf * uz + (Gz - sfn * uz) * ir;
}

// sum_k f_mu(k) * u_x^a u_y^b u_z^c over the valid slots for mu < NMU
// (powers px[k], py[k], pz[k]; f[k * R + mu]).  Slots go in batches
// whose LDS reads are all issued before their arithmetic (the scheduling
// barrier keeps the compiler from interleaving each read with its FMAs,
// which would expose one LDS latency per read).
template <int NMU>
__device__ __forceinline__ void basic_moment_sum(const double* px, const double* py, const double* pz,
                                                 const double* f, int R, int nv, double* acc)
{
    constexpr int UB = NMU <= 2 ? 8 : (NMU <= 4 ? 4 : 2);
    int k = 0;
    for (; k + UB <= nv; k += UB) {
        double a[UB], b[UB], c[UB], fr[UB][NMU];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            a[u] = px[k + u]; b[u] = py[k + u]; c[u] = pz[k + u];
#pragma unroll
            for (int mu = 0; mu < NMU; ++mu) fr[u][mu] = f[(k + u) * R + mu];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const double Pm = a[u] * b[u] * c[u];
#pragma unroll
            for (int mu = 0; mu < NMU; ++mu) acc[mu] = fma(fr[u][mu], Pm, acc[mu]);
        }
    }
    for (; k < nv; ++k) {
        const double Pm = px[k] * py[k] * pz[k];
        const double* fr = f + k * R;
#pragma unroll
        for (int mu = 0; mu < NMU; ++mu) acc[mu] = fma(fr[mu], Pm, acc[mu]);
    }
}

// Basic moments of one row from its power table pw (row stride Kp, np1
// orders per component) and radial values f: monomial t on lane t % 64 of
// wavefront (t / 64) % NW, all its mu; m[ms * x] = value of basic moment x.
// Called by all NW * 64 threads of the row (wave w, lane); ends with a
// barrier.
__device__ __forceinline__ void basic_mono_unit(const DevPlan& P, int t, int k0, int k1, const double* pw, int Kp,
                                                int np1, const double* f, double* acc)
{
    const I4 mo = P.mono[t];
    const int gmax = __builtin_amdgcn_readfirstlane(mo.w >> 24);
    const double* px = pw + (size_t)mo.x * Kp + k0;
    const double* py = pw + (size_t)(np1 + mo.y) * Kp + k0;
    const double* pz = pw + (size_t)(2 * np1 + mo.z) * Kp + k0;
    const double* fk = f + k0 * P.R;
    const int n = k1 - k0;
#pragma unroll
    for (int mu = 0; mu < kMaxR; ++mu) acc[mu] = 0.0;
    switch (gmax) {
        case 0: basic_moment_sum<1>(px, py, pz, fk, P.R, n, acc); break;
        case 1: basic_moment_sum<2>(px, py, pz, fk, P.R, n, acc); break;
        case 2: basic_moment_sum<3>(px, py, pz, fk, P.R, n, acc); break;
        case 3: basic_moment_sum<4>(px, py, pz, fk, P.R, n, acc); break;
        case 4: basic_moment_sum<5>(px, py, pz, fk, P.R, n, acc); break;
        case 5: basic_moment_sum<6>(px, py, pz, fk, P.R, n, acc); break;
        case 6: basic_moment_sum<7>(px, py, pz, fk, P.R, n, acc); break;
        default: basic_moment_sum<8>(px, py, pz, fk, P.R, n, acc); break;
    }
}

__device__ __forceinline__ void store_basic(const DevPlan& P, int t, const double* acc, double* m, int ms)
{
    const I4 d0 = P.mono_dst[2 * t], d1 = P.mono_dst[2 * t + 1];
    const int dst[kMaxR] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
    for (int mu = 0; mu < kMaxR; ++mu)
        if (dst[mu] >= 0) m[ms * dst[mu]] = acc[mu];
}

// The same over slot pairs with 128-bit LDS reads: px, py, pz hold the
// powers of slots (k, k + 1) in one 16-byte pair, f[k * Rp + mu] the radial
// values with mu pairs in 16 bytes (the same address on every lane of the
// wavefront), and nvp (even) slots, the last one a zero slot when the row's
// count is odd.
template <int NMU>
__device__ __forceinline__ void basic_moment_sum2(const double* px, const double* py, const double* pz,
                                                  const double* f, int Rp, int nvp, double* acc)
{
    constexpr int NP = (NMU + 1) / 2;                       // mu pairs
    constexpr int UB = NMU <= 2 ? 4 : (NMU <= 4 ? 2 : 1);   // slot pairs per batch
    int k = 0;
    for (; k + 2 * UB <= nvp; k += 2 * UB) {
        double2 a[UB], b[UB], c[UB], f0[UB][NP], f1[UB][NP];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int kk = k + 2 * u;
            a[u] = *reinterpret_cast<const double2*>(px + kk);
            b[u] = *reinterpret_cast<const double2*>(py + kk);
            c[u] = *reinterpret_cast<const double2*>(pz + kk);
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                f0[u][j] = *reinterpret_cast<const double2*>(f + (size_t)kk * Rp + 2 * j);
                f1[u][j] = *reinterpret_cast<const double2*>(f + (size_t)(kk + 1) * Rp + 2 * j);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const double P0 = a[u].x * b[u].x * c[u].x;
            const double P1 = a[u].y * b[u].y * c[u].y;
#pragma unroll
            for (int mu = 0; mu < NMU; ++mu) {
                const double g0 = (mu & 1) ? f0[u][mu >> 1].y : f0[u][mu >> 1].x;
                const double g1 = (mu & 1) ? f1[u][mu >> 1].y : f1[u][mu >> 1].x;
                acc[mu] = fma(g1, P1, fma(g0, P0, acc[mu]));
            }
        }
    }
    for (; k < nvp; k += 2) {
        const double2 a = *reinterpret_cast<const double2*>(px + k);
        const double2 b = *reinterpret_cast<const double2*>(py + k);
        const double2 c = *reinterpret_cast<const double2*>(pz + k);
        const double P0 = a.x * b.x * c.x, P1 = a.y * b.y * c.y;
#pragma unroll
        for (int mu = 0; mu < NMU; ++mu) acc[mu] = fma(f[(size_t)(k + 1) * Rp + mu], P1, fma(f[(size_t)k * Rp + mu], P0, acc[mu]));
    }
}

// Basic moments of one row of the single-row kernels (paired-slot layout of
// basic_moment_sum2; power table row stride Kp2 = pw_stride(K)).
template <int NW>
__device__ __forceinline__ void basic_moments2(const DevPlan& P, int w, int lane, int nvp, const double* pw, int Kp2,
                                               int np1, const double* f, int Rp, double* m)
{
    const int ngrp = (P.n_mono + kWave - 1) / kWave;
    double acc[kMaxR];
    for (int g = w; g < ngrp; g += NW) {
        const int t = g * kWave + lane;
        if (t < P.n_mono) {
            const I4 mo = P.mono[t];
            const int gmax = __builtin_amdgcn_readfirstlane(mo.w >> 24);
            const double* px = pw + (size_t)mo.x * Kp2;
            const double* py = pw + (size_t)(np1 + mo.y) * Kp2;
            const double* pz = pw + (size_t)(2 * np1 + mo.z) * Kp2;
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu) acc[mu] = 0.0;
            switch (gmax) {
                case 0: basic_moment_sum2<1>(px, py, pz, f, Rp, nvp, acc); break;
                case 1: basic_moment_sum2<2>(px, py, pz, f, Rp, nvp, acc); break;
                case 2: basic_moment_sum2<3>(px, py, pz, f, Rp, nvp, acc); break;
                case 3: basic_moment_sum2<4>(px, py, pz, f, Rp, nvp, acc); break;
                case 4: basic_moment_sum2<5>(px, py, pz, f, Rp, nvp, acc); break;
                case 5: basic_moment_sum2<6>(px, py, pz, f, Rp, nvp, acc); break;
                case 6: basic_moment_sum2<7>(px, py, pz, f, Rp, nvp, acc); break;
                default: basic_moment_sum2<8>(px, py, pz, f, Rp, nvp, acc); break;
            }
            store_basic(P, t, acc, m, 1);
        }
    }
    __syncthreads();
}


// Basic moments by monomial columns (the plan's ctask table): a task holds
// the monomials (a, b, c0 + i), i < C, each with NMU mu, and runs on the
// four lanes of a quad; lane s of the quad sums the slot pairs p = s, s + 4,
// ...  Per slot pair it reads u_x^a, u_y^b, u_z^c0 and u_z (16-byte slot
// pairs of the power table) and the radial values once, and walks the
// column by a running product: a column of C monomials costs the LDS reads
// of one.  The four partial sums are added over the quad by DPP moves
// (every lane of the quad then holds the same bits) and lane 0 stores them.
template <int NMU>
struct ColPair {
    double2 X, Y, Z, U;
    double f0[NMU], f1[NMU];
};

template <int NMU, int C>
__device__ __forceinline__ void column_load(ColPair<NMU>& v, int k, const double* px, const double* py, const double* pz,
                                            const double* pu, const double* f, int Rp)
{
    v.X = *reinterpret_cast<const double2*>(px + k);
    v.Y = *reinterpret_cast<const double2*>(py + k);
    v.Z = *reinterpret_cast<const double2*>(pz + k);
    if constexpr (C > 1) v.U = *reinterpret_cast<const double2*>(pu + k);
    if constexpr (NMU == 1) {
        v.f0[0] = f[(size_t)k * Rp];
        v.f1[0] = f[(size_t)(k + 1) * Rp];
    } else {
#pragma unroll
        for (int j = 0; j < NMU / 2; ++j) {
            const double2 a0 = *reinterpret_cast<const double2*>(f + (size_t)k * Rp + 2 * j);
            const double2 a1 = *reinterpret_cast<const double2*>(f + (size_t)(k + 1) * Rp + 2 * j);
            v.f0[2 * j] = a0.x; v.f0[2 * j + 1] = a0.y;
            v.f1[2 * j] = a1.x; v.f1[2 * j + 1] = a1.y;
        }
    }
}

template <int NMU, int C>
__device__ __forceinline__ void column_add(const ColPair<NMU>& v, double* acc)
{
    double q0 = v.X.x * v.Y.x * v.Z.x, q1 = v.X.y * v.Y.y * v.Z.y;
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int mu = 0; mu < NMU; ++mu)
            acc[c * NMU + mu] = fma(v.f1[mu], q1, fma(v.f0[mu], q0, acc[c * NMU + mu]));
        if (c + 1 < C) { q0 *= v.U.x; q1 *= v.U.y; }
    }
}

template <int NMU, int C>
__device__ __forceinline__ void column_task(const I4* task, int s, int npairs, const double* pw, int Kp2, int np1,
                                            const double* f, int Rp, double* m)
{
    const I4 hd = task[0];
    const double* px = pw + (size_t)hd.x * Kp2;
    const double* py = pw + (size_t)(np1 + hd.y) * Kp2;
    const double* pz = pw + (size_t)(2 * np1 + hd.z) * Kp2;
    const double* pu = pw + (size_t)(2 * np1 + 1) * Kp2;
    double acc[NMU * C];
#pragma unroll
    for (int i = 0; i < NMU * C; ++i) acc[i] = 0.0;
    for (int p = s; p < npairs; p += 4) {
        ColPair<NMU> v0;
        column_load<NMU, C>(v0, 2 * p, px, py, pz, pu, f, Rp);
        column_add<NMU, C>(v0, acc);
    }
    const I4 d0 = task[1], d1 = task[2];
    const int dst[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
    for (int i = 0; i < NMU * C; ++i) {
        double v = acc[i];
        v += dpp_partner<0xB1>(v);  // quad_perm [1, 0, 3, 2]
        v += dpp_partner<0x4E>(v);  // quad_perm [2, 3, 0, 1]
        if (s == 0 && dst[i] >= 0) m[dst[i]] = v;
    }
}

// The basic moments of one row by monomial columns (see column_task); the
// power table and f as for basic_moments2.  Called by all NW * 64 threads
// of the row; ends with a barrier.
template <int NW>
__device__ __forceinline__ void basic_moments_cols(const DevPlan& P, int w, int lane, int nvp, const double* pw,
                                                   int Kp2, int np1, const double* f, int Rp, double* m)
{
    const int npairs = nvp >> 1, s = lane & 3;
    for (int g = w; g < P.n_cwaves; g += NW) {  // wave-uniform: every lane active below
        const I4* task = P.ctask + ((size_t)g * kColTasksPerWave + (lane >> 2)) * 3;
        const int shape = __builtin_amdgcn_readfirstlane(task[0].w);
        switch (shape) {
            case 0: column_task<1, 8>(task, s, npairs, pw, Kp2, np1, f, Rp, m); break;
            case 1: column_task<2, 4>(task, s, npairs, pw, Kp2, np1, f, Rp, m); break;
            case 2: column_task<4, 2>(task, s, npairs, pw, Kp2, np1, f, Rp, m); break;
            case 3: column_task<6, 1>(task, s, npairs, pw, Kp2, np1, f, Rp, m); break;
            default: column_task<8, 1>(task, s, npairs, pw, Kp2, np1, f, Rp, m); break;
        }
    }
    __syncthreads();
}
static_assert(kColShapes == 5 && kColNmu[0] == 1 && kColLen[0] == 8 && kColNmu[1] == 2 && kColLen[1] == 4 &&
                  kColNmu[2] == 4 && kColLen[2] == 2 && kColNmu[3] == 6 && kColLen[3] == 1 && kColNmu[4] == 8 &&
                  kColLen[4] == 1,
              "basic_moments_cols dispatches these shapes");

template <int NW>
__device__ __forceinline__ void basic_moments(const DevPlan& P, int w, int lane, int nv, const double* pw, int Kp,
                                              int np1, const double* f, double* m, int ms)
{
    const int ngrp = (P.n_mono + kWave - 1) / kWave;
    double acc[kMaxR];
    for (int g = w; g < ngrp; g += NW) {
        const int t = g * kWave + lane;
        if (t < P.n_mono) {
            basic_mono_unit(P, t, 0, nv, pw, Kp, np1, f, acc);
            store_basic(P, t, acc, m, ms);
        }
    }
    __syncthreads();
}

template <typename RadialReal, int NT>
__device__ __forceinline__ void atom_row_v2(const DevPlan& P, const AtomArgs& A, const int row, double* lds)
{
    if (A.skip_row && A.skip_row[row]) return;  // the whole workgroup: no barrier reached yet
    const int tid = threadIdx.x;
    const int K = A.nslots;
    constexpr int NW = NT / kWave;
    const AtomLds2 L = atom_lds2_layout(P.amc, P.n_live, P.B, P.k_layout, P.R, P.nu_max, P.n_parts, NT, P.max_partials,
                                        A.xi_atom ? P.amc : 0, P.lds_tables != 0, P.coef_lds);
    if (tid < kLdsM) lds[tid] = 0.0;  // operand of the padding entries, sink of idle lanes
    double* m = lds + L.m;
    double* dm = lds + L.dm;
    double* part = lds + L.part;
    double* f = lds + L.f;
    double* vx = lds + L.vx;
    double* vy = lds + L.vy;
    double* vz = lds + L.vz;
    double* rr = lds + L.r;
    double* red = lds + L.red;      // [NT] the site energy's per-thread terms (end of the adjoint region)
    double* epart = lds + L.epart;  // [NW] per-wave sums of the site energy
    int* ints = reinterpret_cast<int*>(lds + L.ints);
    int* nbr_t = ints;  // [K] compact index -> neighbour species (the radial basis's only input from the slot)
    int* cidx = ints + K;
    int* wave_cnt = ints + 2 * K;
    int* nvp = wave_cnt + NW;

    const int ti = A.itypes[row];
    const double* rij_row = A.rij + (size_t)row * K * 3;
    const int* jt_row = A.jtypes + (size_t)row * K;
    // radial coefficients of this row's species, [tk][mu][n]; visible after
    // the compaction's barriers
    double* coef = lds + L.coef;
    if (P.coef_lds) {
        const double* src = P.radial_coeffs + (size_t)wrap_species(ti, P.Sp) * P.coef_lds;
        for (int x = tid; x < P.coef_lds; x += NT) coef[x] = src[x];
    }
    if (NT <= 256 && A.mom_in) {  // cached moments; visible after the compaction's barriers
        const double* mi = A.mom_in + (size_t)row * P.n_live;
        for (int x = tid; x < P.n_live; x += NT) m[x] = mi[x];
    }

    // ---- A: compaction of valid slots in slot order (radial basis below)
    int nv = 0;
    if constexpr (NT < kAtomBlock) {  // slots in chunks of NT (K may exceed NT)
        // every chunk's loads issue before the first compaction barrier
        constexpr int MC = kAtomBlock / NT;
        double c0[MC], c1[MC], c2[MC];
        int ct[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            const int k = c * NT + tid;
            c0[c] = 0.0; c1[c] = 0.0; c2[c] = 0.0; ct[c] = -1;
            if (k < K) {
                c0[c] = rij_row[3 * k]; c1[c] = rij_row[3 * k + 1]; c2[c] = rij_row[3 * k + 2];
                ct[c] = jt_row[k];
            }
        }
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            const int k0 = c * NT;
            if (k0 >= K) break;
            const int k = k0 + tid;
            const double v0 = c0[c], v1 = c1[c], v2 = c2[c];
            double r = 1.0;
            bool valid = false;
            if (k < K) {
                r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                valid = (ct[c] >= 0) && (r < P.max_dist);
            }
            flag_nonfinite(A.status, k < K && !isfinite(r));
            const uint64_t mask = __ballot(valid);
            const int lane = tid & (kWave - 1), wave = tid / kWave;
            if (lane == 0) wave_cnt[wave] = __popcll(mask);
            __syncthreads();
            int off = nv;
            for (int w = 0; w < wave; ++w) off += wave_cnt[w];
            const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
            if (k < K) cidx[k] = valid ? pos : -1;
            if (valid) {
                vx[pos] = v0; vy[pos] = v1; vz[pos] = v2; rr[pos] = r;
                nbr_t[pos] = ct[c];
            }
            for (int w = 0; w < NW; ++w) nv += wave_cnt[w];
            nv = __builtin_amdgcn_readfirstlane(nv);
            __syncthreads();
        }
    } else {  // K <= NT: one slot per thread
        const int k = tid;
        double v0 = 0, v1 = 0, v2 = 0, r = 1.0;
        int tk = -1;
        bool valid = false;
        if (k < K) {
            v0 = rij_row[3 * k]; v1 = rij_row[3 * k + 1]; v2 = rij_row[3 * k + 2];
            tk = jt_row[k];
            r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
            valid = (tk >= 0) && (r < P.max_dist);
        }
        flag_nonfinite(A.status, !isfinite(r));
        const uint64_t mask = __ballot(valid);
        const int lane = tid & (kWave - 1), wave = tid / kWave;
        if (lane == 0) wave_cnt[wave] = __popcll(mask);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
        if (k < K) cidx[k] = valid ? pos : -1;
        if (valid) {
            vx[pos] = v0; vy[pos] = v1; vz[pos] = v2; rr[pos] = r;
            nbr_t[pos] = tk;
        }
        if (tid == 0) {
            int tot = 0;
            for (int w = 0; w < NW; ++w) tot += wave_cnt[w];
            *nvp = tot;
        }
        __syncthreads();
        nv = __builtin_amdgcn_readfirstlane(*nvp);
    }

    // ---- B: basic moments m[mu, alpha] = sum_k f_mu(k) u_k^alpha.  The
    // power table pw[(q * np1 + p) * Kp + k] = u_{k,q}^p (slots fastest, odd
    // row stride Kp) and f[k * R + mu] sit in the dm region.  One
    // thread per monomial; the mu range is the largest of the monomial's
    // wavefront, so the slot loop is a wave-uniform unrolled instance (lanes
    // of lower degree carry unused mu).
    if (NT <= 256 && A.mom_in) {
        // moments [0, n_live) of this row as the descriptor pass left them
        // (copied before the compaction): no radial basis, basic moments,
        // forward layers or energy here.  (The 512-thread kernel recomputes
        // them: in it this branch costs registers; the results are the
        // same.)
    } else {
        const int np1 = P.nu_max + 1;
        const int Kp2 = pw_stride(P.k_layout);
        const int Rp = f_stride(P.R);
        const int nvp = (nv + 1) & ~1;  // slots in pairs; an odd count gets a zero slot
        double* pw = dm;
        // radial values f[k * Rp + mu]: one thread per (slot, mu); zero for
        // the padding (mu >= R, the pair's zero slot)
        const unsigned rq = row_quotient_magic(Rp);
        if (P.coef_lds) {
            for (int i = tid; i < nvp * Rp; i += NT) {
                const int c = row_quotient(i, rq), mu = i - c * Rp;
                double fv = 0.0, dfv;
                if (c < nv && mu < P.R)
                    radial_one_c<RadialReal>(P, rr[c], coef + ((size_t)wrap_species(nbr_t[c], P.Sp) * P.R + mu) * P.N,
                                             fv, dfv);
                f[i] = fv;
            }
        } else {
            for (int i = tid; i < nvp * Rp; i += NT) {
                const int c = row_quotient(i, rq), mu = i - c * Rp;
                double fv = 0.0, dfv;
                if (c < nv && mu < P.R) radial_one<RadialReal>(P, rr[c], ti, nbr_t[c], mu, fv, dfv);
                f[i] = fv;
            }
        }
        for (int t = tid; t < nvp * 3; t += NT) {
            const int k = t / 3, q = t - 3 * k;
            const double u = k < nv ? (q == 0 ? vx[k] : (q == 1 ? vy[k] : vz[k])) * (1.0 / rr[k]) : 0.0;
            double* col = pw + (size_t)q * np1 * Kp2 + k;
            double v = 1.0;
            col[0] = 1.0;
            for (int p = 1; p < np1; ++p) { v *= u; col[(size_t)p * Kp2] = v; }
        }
        __syncthreads();
        // by monomial columns where the plan has them (the 256- and
        // 512-thread kernels: level 24 2.29 -> 2.23 ms; the 64- and
        // 128-thread kernels at levels 12-20 are faster per monomial)
        if (P.n_cwaves > 0) basic_moments_cols<NW>(P, tid / kWave, tid & (kWave - 1), nvp, pw, Kp2, np1, f, Rp, m);
        else basic_moments2<NW>(P, tid / kWave, tid & (kWave - 1), nvp, pw, Kp2, np1, f, Rp, m);
        // moments that are never targets stay 0 (m[x >= n_live] lies in the dm
        // region, free once the power table and f are consumed); in every
        // level the reference ships, each non-basic moment is a target
        if (P.zero_untargeted)
            for (int x = P.B + tid; x < P.amc; x += NT) m[x] = 0.0;
        __syncthreads();

        // ---- C: contraction layers (depth 1 .. n_depth)
        // The contraction layers issue the kernel's LDS traffic: their waves
        // take issue priority over those of the other workgroup on the CU
        // that are in the radial, basic-moment or gradient phases (measured:
        // 2.44 -> 2.38 ms at level 24; raising the basic moments as well
        // loses it again).
        __builtin_amdgcn_s_setprio(2);
        for (int d = 0; d < P.n_depth; ++d) contraction_layer<NT, false>(P, d, lds, dm, part, tid, P.xi_of);

        // ---- D: site energy (and the basis values m[mapping] when asked for)
        {
            const double* xrow = A.xi_atom ? A.xi_atom + (size_t)row * A.xi_stride : P.moment_coeffs;
            double acc = 0.0;
            for (int s = tid; s < P.S; s += NT) {
                const double b = m[P.mapping[s]];
                acc += xrow[s] * b;
                if (A.basis) A.basis[(size_t)row * P.S + s] = b;
            }
            if (A.mom_out)
                for (int x = tid; x < P.n_live; x += NT) A.mom_out[(size_t)row * P.n_live + x] = m[x];
            // Per-wave sums now (eight lanes add eight values each, then lane
            // order), their total at the end of the row (in wave order).  The
            // wave reads only its own red[] entries: a wavefront's LDS
            // accesses complete in issue order.  (Lane shuffles here would
            // share their index registers with the virial's and spill them.)
            red[tid] = acc;
            __builtin_amdgcn_wave_barrier();
            const int lane = tid & (kWave - 1);
            double s8 = 0.0;
            if (lane < 8)
                for (int j = 0; j < 8; ++j) s8 += red[(tid - lane) + lane * 8 + j];
            double sw = 0.0;
#pragma unroll
            for (int l = 0; l < 8; ++l) sw += lane_value(s8, l);
            if (lane == 0) epart[tid / kWave] = sw;
            // the reverse pass overwrites the never-operand moments read above
            __syncthreads();
        }
    }
    const bool energy = A.Ei && !(NT <= 256 && A.mom_in);
    auto store_energy = [&]() {
        if (energy && tid == 0) {
            double e = 0.0;
            for (int w = 0; w < NW; ++w) e += epart[w];
            A.Ei[row] = P.species_coeffs[wrap_species(ti, P.Sp)] + e;
        }
    };
    if (A.fwd_only) {
        store_energy();
        return;
    }

    // ---- E: reverse layers, deepest first.  dE/dm starts from xi_of (sum of
    // the coefficients mapped to each moment), or from this row's own
    // coefficient vector summed the same way (distilled-head forces, basis
    // Jacobians).
    const double* xi = P.xi_of;
    if (A.xi_atom) {
        const double* xrow = A.xi_atom + (size_t)row * A.xi_stride;
        double* xl = lds + L.xl;
        for (int x = tid; x < P.amc; x += NT) xl[x] = 0.0;
        __syncthreads();
        for (int q = tid; q < P.n_mapped; q += NT) {
            double v = 0.0;
            for (int t = P.map_beg[q]; t < P.map_beg[q + 1]; ++t) v += xrow[P.map_s[t]];
            xl[P.map_x[q]] = v;
        }
        __syncthreads();
        xi = xl;
    }
    __builtin_amdgcn_s_setprio(2);
    for (int li = 0; li <= P.n_depth; ++li) contraction_layer<NT, true>(P, li, lds, dm, part, tid, xi);
    __builtin_amdgcn_s_setprio(0);

    // ---- F: per-slot gradient, thread = (slot, part).  F reads the
    // adjoints of the basic moments only; the rest of the adjoint region
    // holds the per-part gradients and f, f' recomputed for the valid slots.
    double* gpart = lds + L.gp;
    double* gf = lds + L.gf;
    double* gfp = lds + L.gfp;
    const unsigned rq = row_quotient_magic(P.R);
    if (P.coef_lds) {
        for (int i = tid; i < nv * P.R; i += NT) {
            const int c = row_quotient(i, rq), mu = i - c * P.R;
            double fv, dfv;
            radial_one_c<RadialReal>(P, rr[c], coef + ((size_t)wrap_species(nbr_t[c], P.Sp) * P.R + mu) * P.N, fv, dfv);
            gf[mu * K + c] = fv;
            gfp[mu * K + c] = dfv;
        }
    } else {
        for (int i = tid; i < nv * P.R; i += NT) {
            const int c = row_quotient(i, rq), mu = i - c * P.R;
            double fv, dfv;
            radial_one<RadialReal>(P, rr[c], ti, nbr_t[c], mu, fv, dfv);
            gf[mu * K + c] = fv;
            gfp[mu * K + c] = dfv;
        }
    }
    __syncthreads();
    {
        const int kk = tid & (kWave - 1), prt = tid / kWave;
        const int nu_max = P.nu_max;
        for (int k = kk; k < nv && prt < P.n_parts; k += kWave) {
            if (nu_max <= 4) slot_gradient<4>(P, prt, k, K, dm, gf, gfp, vx, vy, vz, rr, gpart);
            else if (nu_max <= 6) slot_gradient<6>(P, prt, k, K, dm, gf, gfp, vx, vy, vz, rr, gpart);
            else if (nu_max <= 8) slot_gradient<8>(P, prt, k, K, dm, gf, gfp, vx, vy, vz, rr, gpart);
            else if (nu_max <= 10) slot_gradient<10>(P, prt, k, K, dm, gf, gfp, vx, vy, vz, rr, gpart);
            else slot_gradient<kMaxNu>(P, prt, k, K, dm, gf, gfp, vx, vy, vz, rr, gpart);
        }
    }
    __syncthreads();
    double* gk = dm;
    for (int t = tid; t < nv * 3; t += NT) {
        double s = 0.0;
        for (int prt = 0; prt < P.n_parts; ++prt) s += gpart[(size_t)prt * K * 3 + t];
        gk[t] = s;
    }
    __syncthreads();
    double* g_row = A.g + (size_t)row * K * 3;
    for (int t = tid; t < K * 3; t += NT) {
        const int k = t / 3, q = t - 3 * k;
        const int c = cidx[k];
        g_row[t] = c >= 0 ? gk[3 * c + q] : 0.0;
    }
    // ---- G: per-atom virial, one component per wavefront (lanes over slots)
    if (A.Wi) {
        const int lane = tid & (kWave - 1);
        for (int c = tid / kWave; c < 9; c += NW) {
            const int a = c / 3, b = c - 3 * (c / 3);
            const double* va = a == 0 ? vx : (a == 1 ? vy : vz);
            double s = 0.0;
            for (int k = lane; k < nv; k += kWave) s += va[k] * gk[3 * k + b];
            s = wave_sum(s);
            if (lane == 0) A.Wi[(size_t)row * 9 + c] = s;
        }
    }
    store_energy();
}

// One workgroup per row.  512-thread workgroups are compiled for two per CU
// (<= 128 VGPRs; the LDS layout keeps level 24 under half of the CU's LDS).
template <typename RadialReal, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 512 ? 4 : 1)))
atom_kernel_v2(const DevPlan P, AtomArgs A)
{
    extern __shared__ double lds[];
    atom_row_v2<RadialReal, NT>(P, A, blockIdx.x, lds);
}

// The rows of a list, walked by a fixed grid (the list length is read on the
// device, so the list can come from a device-side compaction without a host
// round trip).  A separate kernel: the loop costs the per-row kernel ~40
// VGPRs and spills.
template <int NT>
__global__ void __launch_bounds__(NT) atom_kernel_v2_list(const DevPlan P, AtomArgs A)
{
    extern __shared__ double lds[];
    const int n = *A.row_count;
    for (int idx = blockIdx.x; idx < n; idx += gridDim.x) {
        atom_row_v2<double, NT>(P, A, A.rows[idx], lds);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Paired rows: one workgroup evaluates two atoms.  Their moments, adjoints
// and partial sums are interleaved in 16-byte LDS slots, so each entry of a
// contraction stream (its table load, operand offsets and FP32 multiplier)
// serves both rows with one 128-bit LDS read per operand.  The per-atom
// phases (radial basis, basic moments, energy, slot gradients, virial) run
// on one half of the workgroup per row, exactly as the single-row kernel.
// Operand fields are slot indices from the LDS base (StreamLayout absolute).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stream_steps_pair(const uint4* q, char* lb, double2& acc)
{
    constexpr int U = kStreamUnroll;
    double2 a0[U], b0[U], a1[U], b1[U], a2[U], b2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        a0[u] = *reinterpret_cast<const double2*>(lb + ((q[u].x & 0xffffu) << 4));
        b0[u] = *reinterpret_cast<const double2*>(lb + ((q[u].x >> 16) << 4));
        a1[u] = *reinterpret_cast<const double2*>(lb + ((q[u].y & 0xffffu) << 4));
        b1[u] = *reinterpret_cast<const double2*>(lb + ((q[u].y >> 16) << 4));
        a2[u] = *reinterpret_cast<const double2*>(lb + ((q[u].z & 0xffffu) << 4));
        b2[u] = *reinterpret_cast<const double2*>(lb + ((q[u].z >> 16) << 4));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        // three entries with one integer multiplier; on the last step of
        // the lane's item (per lane) the sums go to the item's destination
        // slot and restart from zero
        double2 t;
        t.x = a0[u].x * b0[u].x;
        t.y = a0[u].y * b0[u].y;
        t.x = fma(a1[u].x, b1[u].x, t.x);
        t.y = fma(a1[u].y, b1[u].y, t.y);
        t.x = fma(a2[u].x, b2[u].x, t.x);
        t.y = fma(a2[u].y, b2[u].y, t.y);
        const uint32_t w = q[u].w;
        const double m = (double)stream_mult(w);
        acc.x = fma(m, t.x, acc.x);
        acc.y = fma(m, t.y, acc.y);
        const uint32_t dst = stream_dest(w);
        if (dst) {
            *reinterpret_cast<double2*>(lb + dst) = acc;
            acc = make_double2(0.0, 0.0);
        }
    }
}

template <int NT, bool REV>
__device__ __forceinline__ void contraction_layer_pair(const DevPlan& P, int L, double* lds, double* m2, double* dm2,
                                                       double* part2, int tid)
{
    constexpr int NW = NT / kWave;
    constexpr int U = kStreamUnroll;
    const int lane = tid & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
    const I4 wv = (REV ? P.rev_wave : P.fwd_wave)[L * NW + w];
    const int nsteps = wv.y;
    if (nsteps > 0) {
        const uint4* tb = reinterpret_cast<const uint4*>(REV ? P.rev_tbl : P.fwd_tbl) + (size_t)wv.x * kWave + lane;
        char* lb = reinterpret_cast<char*>(lds);
        double2 acc = make_double2(0.0, 0.0);
        uint4 qa[U], qb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) qa[u] = tb[(size_t)u * kWave];
        for (int s0 = 0;;) {
            if (s0 + U < nsteps) {
#pragma unroll
                for (int u = 0; u < U; ++u) qb[u] = tb[(size_t)(s0 + U + u) * kWave];
            }
            stream_steps_pair(qa, lb, acc);
            s0 += U;
            if (s0 >= nsteps) break;
            if (s0 + U < nsteps) {
#pragma unroll
                for (int u = 0; u < U; ++u) qa[u] = tb[(size_t)(s0 + U + u) * kWave];
            }
            stream_steps_pair(qb, lb, acc);
            s0 += U;
            if (s0 >= nsteps) break;
        }
    }
    __syncthreads();
    const int c0 = REV ? P.rev_cb_beg[L] : P.fwd_cb_beg[L];
    const int c1 = REV ? P.rev_cb_beg[L + 1] : P.fwd_cb_beg[L + 1];
    const I4* comb = REV ? P.rev_comb : P.fwd_comb;
    double2* dst = reinterpret_cast<double2*>(REV ? dm2 : m2);
    const double2* part = reinterpret_cast<const double2*>(part2);
    for (int c = c0 + tid; c < c1; c += NT) {
        const I4 cb = comb[c];
        const double x0 = REV ? P.xi_of[cb.x] : 0.0;
        double2 acc = make_double2(x0, x0);
        for (int p = cb.y; p < cb.z; ++p) {
            const double2 v = part[p];
            acc.x += v.x;
            acc.y += v.y;
        }
        dst[cb.x] = acc;
    }
    if (REV) {
        for (int t = P.rev_bo_beg[L] + tid; t < P.rev_bo_beg[L + 1]; t += NT) {
            const int x = P.rev_bo[t];
            dst[x] = make_double2(P.xi_of[x], P.xi_of[x]);
        }
        for (int t = P.rev_xa_beg[L] + tid; t < P.rev_xa_beg[L + 1]; t += NT) {
            const int x = P.rev_xa[t];
            const double v = P.xi_of[x];
            dst[x].x += v;
            dst[x].y += v;
        }
    }
    if (c1 > c0 || REV) __syncthreads();
}

template <typename RadialReal, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 1024 ? 4 : 2)))
atom_kernel_pair(const DevPlan P, AtomArgs A)
{
    extern __shared__ double lds[];
    constexpr int NH = NT / 2, NWH = NH / kWave;
    const int tid = threadIdx.x;
    const int h = tid / NH, ht = tid - h * NH;  // row of the pair, thread within the row
    const int lane = tid & (kWave - 1), wr = ht / kWave;
    const int row = 2 * blockIdx.x + h;
    const bool active = row < A.nrows;
    const int K = A.nslots, Kl = P.k_layout;
    const AtomLdsPair L = atom_lds_pair_layout(P.amc, P.n_live, Kl, P.R, P.nu_max, P.n_parts, NT, P.max_partials);
    double* m2 = lds + 2 * L.m;          // m of row r of moment x: m2[2 x + r]
    double* dm2 = lds + 2 * L.dm;
    double* part2 = lds + 2 * L.part;
    double* vx = lds + L.vec + h * 4 * Kl;
    double* vy = vx + Kl;
    double* vz = vy + Kl;
    double* rr = vz + Kl;
    int* ints = reinterpret_cast<int*>(lds + L.ints) + h * L.ints_per_row;
    int* nbr_t = ints;  // [Kl] compact index -> neighbour species
    int* cidx = ints + Kl;
    int* wave_cnt = ints + 2 * Kl;
    int* nvp = wave_cnt + NWH;
    double* scr_dm = dm2 + h * L.dlen;   // this row's scratch in the adjoint region (A/B, gk)
    const int pw_len = (Kl | 1) * 3 * (P.nu_max + 1);
    double* f = scr_dm + pw_len;
    if (tid < 2 * kPairM) lds[tid] = 0.0;  // zero pair: operand of padding entries, sink of idle lanes

    const int ti = active ? A.itypes[row] : 0;
    const double* rij_row = A.rij + (size_t)(active ? row : 0) * K * 3;
    const int* jt_row = A.jtypes + (size_t)(active ? row : 0) * K;

    // ---- A: compaction of valid slots (K <= NH slots per row; radial basis below)
    int nv;
    {
        const int k = ht;
        double v0 = 0, v1 = 0, v2 = 0, r = 1.0;
        int tk = -1;
        bool valid = false;
        if (active && k < K) {
            v0 = rij_row[3 * k]; v1 = rij_row[3 * k + 1]; v2 = rij_row[3 * k + 2];
            tk = jt_row[k];
            r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
            valid = (tk >= 0) && (r < P.max_dist);
        }
        flag_nonfinite(A.status, !isfinite(r));
        const uint64_t mask = __ballot(valid);
        if (lane == 0) wave_cnt[wr] = __popcll(mask);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wr; ++w) off += wave_cnt[w];
        const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
        if (k < K) cidx[k] = valid ? pos : -1;
        if (valid) {
            vx[pos] = v0; vy[pos] = v1; vz[pos] = v2; rr[pos] = r;
            nbr_t[pos] = tk;
        }
        if (ht == 0) {
            int tot = 0;
            for (int w = 0; w < NWH; ++w) tot += wave_cnt[w];
            *nvp = tot;
        }
        __syncthreads();
        nv = __builtin_amdgcn_readfirstlane(*nvp);
    }

    // ---- B: basic moments of this row (as in the single-row kernel), stored
    // interleaved
    {
        const int np1 = P.nu_max + 1;
        const int Kp = Kl | 1;
        double* pw = scr_dm;
        for (int i = ht; i < nv * P.R; i += NH) {  // radial values: one thread per (slot, mu)
            const int c = i / P.R, mu = i - c * P.R;
            double fv, dfv;
            radial_one<RadialReal>(P, rr[c], ti, nbr_t[c], mu, fv, dfv);
            f[i] = fv;
        }
        for (int t = ht; t < nv * 3; t += NH) {
            const int k = t / 3, q = t - 3 * k;
            const double inv = 1.0 / rr[k];
            const double u = (q == 0 ? vx[k] : (q == 1 ? vy[k] : vz[k])) * inv;
            double* col = pw + (size_t)q * np1 * Kp + k;
            double v = 1.0;
            col[0] = 1.0;
            for (int p = 1; p < np1; ++p) { v *= u; col[(size_t)p * Kp] = v; }
        }
        __syncthreads();
        basic_moments<NWH>(P, wr, lane, nv, pw, Kp, np1, f, m2 + h, 2);
        // moments that are never targets stay 0 (both rows)
        for (int x = P.B + tid; x < P.amc; x += NT) reinterpret_cast<double2*>(m2)[x] = make_double2(0.0, 0.0);
        __syncthreads();
    }

    // ---- C: contraction layers, both rows
    for (int d = 0; d < P.n_depth; ++d) contraction_layer_pair<NT, false>(P, d, lds, m2, dm2, part2, tid);

    // ---- D: site energy of this row
    {
        double acc = 0.0;
        for (int s = ht; s < P.S; s += NH) acc += P.moment_coeffs[s] * m2[2 * P.mapping[s] + h];
        acc = wave_sum(acc);
        // after the never-operand moments, which alias the adjoint region
        double* red = dm2 + 2 * (P.amc - P.n_live) + h * NWH;
        if (lane == 0) red[wr] = acc;
        __syncthreads();
        if (ht == 0 && active) {
            double e = 0.0;
            for (int w = 0; w < NWH; ++w) e += red[w];
            A.Ei[row] = P.species_coeffs[wrap_species(ti, P.Sp)] + e;
        }
        __syncthreads();
    }

    // ---- E: reverse layers, deepest first, both rows
    for (int li = 0; li <= P.n_depth; ++li) contraction_layer_pair<NT, true>(P, li, lds, m2, dm2, part2, tid);

    // ---- F: per-slot gradient of this row; its half of the dead moment
    // region holds the per-part gradients and f, f'
    double* gpart = m2 + h * L.mlen;
    double* gf = gpart + P.n_parts * Kl * 3;
    double* gfp = gf + P.R * Kl;
    for (int i = ht; i < nv * P.R; i += NH) {
        const int c = i / P.R, mu = i - c * P.R;
        double fv, dfv;
        radial_one<RadialReal>(P, rr[c], ti, nbr_t[c], mu, fv, dfv);
        gf[mu * Kl + c] = fv;
        gfp[mu * Kl + c] = dfv;
    }
    __syncthreads();
    {
        const int prt = wr;
        const int nu_max = P.nu_max;
        const double* dmr = dm2 + h;
        for (int k = lane; k < nv && prt < P.n_parts; k += kWave) {
            if (nu_max <= 4) slot_gradient<4, 2>(P, prt, k, Kl, dmr, gf, gfp, vx, vy, vz, rr, gpart);
            else if (nu_max <= 6) slot_gradient<6, 2>(P, prt, k, Kl, dmr, gf, gfp, vx, vy, vz, rr, gpart);
            else if (nu_max <= 8) slot_gradient<8, 2>(P, prt, k, Kl, dmr, gf, gfp, vx, vy, vz, rr, gpart);
            else if (nu_max <= 10) slot_gradient<10, 2>(P, prt, k, Kl, dmr, gf, gfp, vx, vy, vz, rr, gpart);
            else slot_gradient<kMaxNu, 2>(P, prt, k, Kl, dmr, gf, gfp, vx, vy, vz, rr, gpart);
        }
    }
    __syncthreads();
    double* gk = scr_dm;  // the adjoints are dead
    for (int t = ht; t < nv * 3; t += NH) {
        double s = 0.0;
        for (int prt = 0; prt < P.n_parts; ++prt) s += gpart[(size_t)prt * Kl * 3 + t];
        gk[t] = s;
    }
    __syncthreads();
    if (!active) return;  // no barrier follows
    double* g_row = A.g + (size_t)row * K * 3;
    for (int t = ht; t < K * 3; t += NH) {
        const int k = t / 3, q = t - 3 * k;
        const int c = cidx[k];
        g_row[t] = c >= 0 ? gk[3 * c + q] : 0.0;
    }
    // ---- G: per-atom virial, one component per wavefront of the row
    if (A.Wi) {
        for (int c = wr; c < 9; c += NWH) {
            const int a = c / 3, b = c - 3 * (c / 3);
            const double* va = a == 0 ? vx : (a == 1 ? vy : vz);
            double s = 0.0;
            for (int k = lane; k < nv; k += kWave) s += va[k] * gk[3 * k + b];
            s = wave_sum(s);
            if (lane == 0) A.Wi[(size_t)row * 9 + c] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Newton-3 force gather (jax_backend.py:263-303) without FP64 atomics:
// a reverse-neighbor CSR is built per call; each atom sums its incoming
// slots in increasing slot order, so results are bitwise reproducible.
// ---------------------------------------------------------------------------
// The reverse-neighbor CSR (below) depends only on js, n_f and the array
// shape.  csr_key_kernel and csr_compare_kernel set state[0] = 1 when all
// of them equal those of the previous build (kept in state[1..4] and a copy
// of js), and the build kernels then return at once: in an MD run js stays
// the same between neighbor-list rebuilds.  No host synchronisation.
__global__ void csr_key_kernel(int* state, int nrows, int nslots, int n_f)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        state[0] = (state[4] == 1 && state[1] == nrows && state[2] == nslots && state[3] == n_f) ? 1 : 0;
        state[1] = nrows; state[2] = nslots; state[3] = n_f; state[4] = 1;
    }
}

__global__ void csr_compare_kernel(const int* js, int* js_copy, long long total, int* state)
{
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < total;
         s += (long long)gridDim.x * blockDim.x) {
        const int v = js[s];
        if (js_copy[s] != v) {
            js_copy[s] = v;
            state[0] = 0;
        }
    }
}

__global__ void zero_ints_kernel(int* p, int n, const int* skip)
{
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void rev_count_kernel(const int* js, int nrows, int nslots, int n_f, int* cnt, const int* skip)
{
    if (skip && *skip) return;
    const long long total = (long long)nrows * nslots;
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < total;
         s += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(s / nslots);
        const int j = js[s];
        if (i < n_f && j >= 0 && j < n_f && j < nrows) atomicAdd(&cnt[j], 1);
    }
}

// Exclusive scan of cnt[0..n) into off[0..n] in three passes over tiles of
// kScanTile counts: tile sums, a scan of the tile sums, then each tile's own
// scan from its offset.
constexpr int kScanTile = 4096;  // 256 threads x 16

__device__ __forceinline__ int block_excl_scan_256(int v, int* sh, int& total)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = 1; w < 256; w <<= 1) {
        const int add = t >= w ? sh[t - w] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int incl = sh[t];
    total = sh[255];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(256) scan_tiles_kernel(const int* cnt, int n, int* tile_sum, const int* skip)
{
    if (skip && *skip) return;
    __shared__ int sh[256];
    const int base = blockIdx.x * kScanTile;
    int s = 0;
    for (int q = 0; q < kScanTile / 256; ++q) {
        const int i = base + q * 256 + threadIdx.x;
        if (i < n) s += cnt[i];
    }
    int total;
    block_excl_scan_256(s, sh, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) scan_top_kernel(const int* tile_sum, int ntiles, int n, int* tile_off, int* off,
                                                       const int* skip)
{
    if (skip && *skip) return;
    __shared__ int sh[256];
    int carry = 0;
    for (int b0 = 0; b0 < ntiles; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const int v = b < ntiles ? tile_sum[b] : 0;
        int total;
        const int ex = block_excl_scan_256(v, sh, total);
        if (b < ntiles) tile_off[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) off[n] = carry;
}

__global__ void __launch_bounds__(256) scan_apply_kernel(const int* cnt, int n, const int* tile_off, int* off,
                                                         const int* skip)
{
    if (skip && *skip) return;
    __shared__ int tile[kScanTile];
    __shared__ int sh[256];
    const int base = blockIdx.x * kScanTile;
    for (int q = 0; q < kScanTile / 256; ++q) {  // coalesced load of the tile
        const int i = base + q * 256 + threadIdx.x;
        tile[q * 256 + threadIdx.x] = i < n ? cnt[i] : 0;
    }
    __syncthreads();
    constexpr int C = kScanTile / 256;  // each thread scans 16 consecutive counts
    int v[C];
    int s = 0;
#pragma unroll
    for (int q = 0; q < C; ++q) { v[q] = tile[threadIdx.x * C + q]; s += v[q]; }
    int total;
    int run = tile_off[blockIdx.x] + block_excl_scan_256(s, sh, total);
#pragma unroll
    for (int q = 0; q < C; ++q) {
        tile[threadIdx.x * C + q] = run;
        run += v[q];
    }
    __syncthreads();
    for (int q = 0; q < kScanTile / 256; ++q) {
        const int i = base + q * 256 + threadIdx.x;
        if (i < n) off[i] = tile[q * 256 + threadIdx.x];
    }
}

__global__ void rev_fill_kernel(const int* js, int nrows, int nslots, int n_f, const int* off, int* cur, int* rev,
                                const int* skip)
{
    if (skip && *skip) return;
    const long long total = (long long)nrows * nslots;
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < total;
         s += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(s / nslots);
        const int j = js[s];
        if (i < n_f && j >= 0 && j < n_f && j < nrows) {
            const int p = atomicAdd(&cur[j], 1);
            rev[off[j] + p] = (int)s;
        }
    }
}

// One wavefront per atom: its own slots strided over the lanes, its incoming
// slots sorted by slot index (64-lane bitonic sort; a selection loop for
// longer lists) so that the summation order does not depend on the order in
// which rev_fill's atomics placed them.  Bitwise reproducible.
__global__ void __launch_bounds__(256) force_kernel(const double* g, int nrows, int nslots, int n_f, const int* off,
                                                    const int* rev, double* F)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int a = (int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave);
    if (a >= nrows) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (a < n_f) {
        const double* ga = g + (size_t)a * nslots * 3;
        for (int k = lane; k < nslots; k += kWave) { fx += ga[3 * k]; fy += ga[3 * k + 1]; fz += ga[3 * k + 2]; }
        const int s0 = off[a], cnt = off[a + 1] - s0;
        if (cnt <= kWave) {
            int key = lane < cnt ? rev[s0 + lane] : 0x7fffffff;
#pragma unroll
            for (int size = 2; size <= kWave; size <<= 1) {
#pragma unroll
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    const int other = __shfl_xor(key, stride);
                    const bool up = (lane & size) == 0;
                    const bool lower = (lane & stride) == 0;
                    key = (lower == up) ? min(key, other) : max(key, other);
                }
            }
            if (lane < cnt) {
                const double* gs = g + (size_t)key * 3;
                fx -= gs[0]; fy -= gs[1]; fz -= gs[2];
            }
        } else {  // long incoming list: ascending selection, accumulated by lane 0
            int last = -1;
            for (int r = 0; r < cnt; ++r) {
                int m = 0x7fffffff;
                for (int q = lane; q < cnt; q += kWave) {
                    const int v = rev[s0 + q];
                    if (v > last && v < m) m = v;
                }
#pragma unroll
                for (int o = kWave / 2; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o));
                last = m;
                if (lane == 0) {
                    const double* gs = g + (size_t)m * 3;
                    fx -= gs[0]; fy -= gs[1]; fz -= gs[2];
                }
            }
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            fx += __shfl_xor(fx, o);
            fy += __shfl_xor(fy, o);
            fz += __shfl_xor(fz, o);
        }
    }
    if (lane == 0) { F[3 * (size_t)a] = fx; F[3 * (size_t)a + 1] = fy; F[3 * (size_t)a + 2] = fz; }
}

// out[0] = E, out[1..6] = stress (reference Voigt xx yy zz yz xz xy, /V),
// out[7..12] = LAMMPS virial (xx yy zz xy xz yz, energy units), out[13] =
// status: bit 1 a non-finite energy or virial, bit 2 a non-finite slot
// vector r_ij (the atom kernels' flag, which this kernel clears; the energy
// is then NaN, as the reference's).
// Stage 1: block b sums rows b, b + G, ... (E_i for i < n_e, and the 9
// virial components) into tot[b][10]; stage 2 adds the kTotalsBlocks partials
// in block order.  Deterministic for a given row count.
__global__ void __launch_bounds__(256) totals_partial_kernel(const double* Ei, const double* Wi, int nrows, int n_e,
                                                             double* tot)
{
    __shared__ double part[256 / kWave][10];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    double s[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) s[q] = 0.0;
    for (int i = blockIdx.x * 256 + tid; i < nrows; i += kTotalsBlocks * 256) {
        if (i < n_e) s[0] += Ei[i];
        const double* w = Wi + (size_t)i * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) s[1 + q] += w[q];
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        double v = s[q];
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) part[wave][q] = v;
    }
    __syncthreads();
    if (tid < 10) {
        double v = 0.0;
        for (int w = 0; w < 256 / kWave; ++w) v += part[w][tid];
        tot[blockIdx.x * 10 + tid] = v;
    }
}

__global__ void __launch_bounds__(64) totals_kernel(const double* tot, int cell_rank, double volume, double* out,
                                                    int* in_status)
{
    __shared__ double acc[10];
    const int tid = threadIdx.x;
    if (tid < 10) {
        double v = 0.0;
        for (int b = 0; b < kTotalsBlocks; ++b) v += tot[b * 10 + tid];
        acc[tid] = v;
    }
    __syncthreads();
    if (tid == 0) {
        const double* W = acc + 1;
        out[0] = acc[0];
        const int idx[6] = {0, 4, 8, 5, 2, 1};
        for (int q = 0; q < 6; ++q) {
            const int a = idx[q] / 3, b = idx[q] % 3;
            out[1 + q] = cell_rank == 3 ? (W[3 * a + b] + W[3 * b + a]) * (0.5 / volume) : __builtin_nan("");
        }
        // LAMMPS virial = -sym(W), order xx yy zz xy xz yz
        out[7] = -W[0];
        out[8] = -W[4];
        out[9] = -W[8];
        out[10] = -0.5 * (W[1] + W[3]);
        out[11] = -0.5 * (W[2] + W[6]);
        out[12] = -0.5 * (W[5] + W[7]);
        const int bad_in = in_status ? *in_status : 0;
        if (bad_in) {
            out[0] = __builtin_nan("");
            *in_status = 0;
        }
        bool bad = !isfinite(out[0]);
        for (int q = 0; q < 9; ++q) bad = bad || !isfinite(W[q]);
        out[13] = (double)((bad ? 1 : 0) | (bad_in ? 2 : 0));
    }
}

// ---------------------------------------------------------------------------
// Raises (never lowers) the dynamic-LDS limit of a variant's kernels: the
// limit belongs to the kernel function, which every open potential shares.
int atom_kernel_blocks_per_cu(int variant, size_t bytes)
{
    const void* fn = nullptr;
    int nt = 0;
    switch (variant) {
        case kVariantV2x256: fn = (const void*)atom_kernel_v2<double, 256>; nt = 256; break;
        case kVariantV2x512: fn = (const void*)atom_kernel_v2<double, 512>; nt = 512; break;
        case kVariantV2x64: fn = (const void*)atom_kernel_v2<double, 64>; nt = 64; break;
        case kVariantV2x128: fn = (const void*)atom_kernel_v2<double, 128>; nt = 128; break;
        default: return -1;
    }
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, nt, bytes) != hipSuccess) return -1;
    return n;
}

// The single-row kernels' contraction streams store by absolute LDS offset
// (lds_store), which holds while their dynamic LDS array starts at address
// 0: refuse to run one that has static LDS.
template <typename Kern>
static hipError_t no_static_lds(Kern kernel)
{
    hipFuncAttributes a;
    const hipError_t e = hipFuncGetAttributes(&a, (const void*)kernel);
    if (e != hipSuccess) return e;
    return a.sharedSizeBytes == 0 ? hipSuccess : hipErrorInvalidConfiguration;
}

template <int NT>
static hipError_t single_row_kernels_checked()
{
    hipError_t e = no_static_lds(atom_kernel_v2<double, NT>);
    if (e == hipSuccess) e = no_static_lds(atom_kernel_v2<float, NT>);
    if (e == hipSuccess) e = no_static_lds(atom_kernel_v2_list<NT>);
    return e;
}

hipError_t set_atom_kernel_lds(int variant, size_t bytes)
{
    static size_t cur[kNumVariants] = {};
    if (variant < 0 || variant >= kNumVariants) return hipErrorInvalidValue;
    if (bytes <= cur[variant]) return hipSuccess;
    if (cur[variant] == 0) {
        hipError_t c = hipSuccess;
        if (variant == kVariantV2x64) c = single_row_kernels_checked<64>();
        else if (variant == kVariantV2x128) c = single_row_kernels_checked<128>();
        else if (variant == kVariantV2x256) c = single_row_kernels_checked<256>();
        else if (variant == kVariantV2x512) c = single_row_kernels_checked<512>();
        if (c != hipSuccess) return c;
    }
    hipError_t e = set_atom_kernel_lds_exact(variant, bytes);
    if (e == hipSuccess) cur[variant] = bytes;
    return e;
}

hipError_t set_atom_kernel_lds_exact(int variant, size_t bytes)
{
    const int b = (int)bytes;
    switch (variant) {
        case kVariantV2x256: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel_v2<double, 256>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)atom_kernel_v2_list<256>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel_v2<float, 256>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
        case kVariantV2x512: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel_v2<double, 512>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)atom_kernel_v2_list<512>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel_v2<float, 512>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
        case kVariantV2x64: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel_v2<double, 64>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)atom_kernel_v2_list<64>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel_v2<float, 64>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
        case kVariantV2x128: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel_v2<double, 128>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)atom_kernel_v2_list<128>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel_v2<float, 128>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
        case kVariantPair1024: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel_pair<double, 1024>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel_pair<float, 1024>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
        case kVariantPair512: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel_pair<double, 512>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel_pair<float, 512>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
        default: {
            hipError_t e = hipFuncSetAttribute((const void*)atom_kernel<double, false>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)atom_kernel<float, false>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void*)atom_kernel<double, true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)atom_kernel<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       b);
        }
    }
}

hipError_t launch_atoms(const DevPlan& P, const AtomArgs& A, size_t lds_bytes, bool radial_fp32, hipStream_t stream,
                        int variant, int grid)
{
    const bool modes = A.xi_atom || A.basis || A.fwd_only || A.rows || A.row_count || !A.Ei;
    if (grid <= 0) return hipSuccess;
    if (A.rows || A.row_count) {
        if (!A.rows || !A.row_count || radial_fp32) return hipErrorInvalidValue;
        if (variant == kVariantV2x256)
            hipLaunchKernelGGL(atom_kernel_v2_list<256>, dim3(grid), dim3(256), lds_bytes, stream, P, A);
        else if (variant == kVariantV2x512)
            hipLaunchKernelGGL(atom_kernel_v2_list<512>, dim3(grid), dim3(512), lds_bytes, stream, P, A);
        else if (variant == kVariantV2x64)
            hipLaunchKernelGGL(atom_kernel_v2_list<64>, dim3(grid), dim3(64), lds_bytes, stream, P, A);
        else if (variant == kVariantV2x128)
            hipLaunchKernelGGL(atom_kernel_v2_list<128>, dim3(grid), dim3(128), lds_bytes, stream, P, A);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (variant_paired(variant)) {  // plain force evaluation only
        if (modes || !A.g) return hipErrorInvalidValue;
        const int pairs = (grid + 1) / 2;
        if (variant == kVariantPair1024) {
            if (radial_fp32) hipLaunchKernelGGL((atom_kernel_pair<float, 1024>), dim3(pairs), dim3(1024), lds_bytes, stream, P, A);
            else hipLaunchKernelGGL((atom_kernel_pair<double, 1024>), dim3(pairs), dim3(1024), lds_bytes, stream, P, A);
        } else {
            if (radial_fp32) hipLaunchKernelGGL((atom_kernel_pair<float, 512>), dim3(pairs), dim3(512), lds_bytes, stream, P, A);
            else hipLaunchKernelGGL((atom_kernel_pair<double, 512>), dim3(pairs), dim3(512), lds_bytes, stream, P, A);
        }
        return hipGetLastError();
    }
    switch (variant) {
        case kVariantV2x256:
            if (radial_fp32) hipLaunchKernelGGL((atom_kernel_v2<float, 256>), dim3(grid), dim3(256), lds_bytes, stream, P, A);
            else hipLaunchKernelGGL((atom_kernel_v2<double, 256>), dim3(grid), dim3(256), lds_bytes, stream, P, A);
            break;
        case kVariantV2x512:
            if (radial_fp32) hipLaunchKernelGGL((atom_kernel_v2<float, 512>), dim3(grid), dim3(512), lds_bytes, stream, P, A);
            else hipLaunchKernelGGL((atom_kernel_v2<double, 512>), dim3(grid), dim3(512), lds_bytes, stream, P, A);
            break;
        case kVariantV2x64:
            if (radial_fp32) hipLaunchKernelGGL((atom_kernel_v2<float, 64>), dim3(grid), dim3(64), lds_bytes, stream, P, A);
            else hipLaunchKernelGGL((atom_kernel_v2<double, 64>), dim3(grid), dim3(64), lds_bytes, stream, P, A);
            break;
        case kVariantV2x128:
            if (radial_fp32) hipLaunchKernelGGL((atom_kernel_v2<float, 128>), dim3(grid), dim3(128), lds_bytes, stream, P, A);
            else hipLaunchKernelGGL((atom_kernel_v2<double, 128>), dim3(grid), dim3(128), lds_bytes, stream, P, A);
            break;
        default:
            if (modes) return hipErrorInvalidValue;  // the per-target kernel has no optional modes
            if (P.mscratch) {  // moments in global memory: a fixed grid walks the rows
                const int g = grid < P.mscratch_grid ? grid : P.mscratch_grid;
                if (radial_fp32) hipLaunchKernelGGL((atom_kernel<float, true>), dim3(g), dim3(kBlock), lds_bytes, stream, P, A);
                else hipLaunchKernelGGL((atom_kernel<double, true>), dim3(g), dim3(kBlock), lds_bytes, stream, P, A);
            } else if (radial_fp32) {
                hipLaunchKernelGGL((atom_kernel<float, false>), dim3(grid), dim3(kBlock), lds_bytes, stream, P, A);
            } else {
                hipLaunchKernelGGL((atom_kernel<double, false>), dim3(grid), dim3(kBlock), lds_bytes, stream, P, A);
            }
            break;
    }
    return hipGetLastError();
}

hipError_t launch_gather(const AtomArgs& A, const ForceArgs& F, hipStream_t stream)
{
    const int n = A.nrows;
    if (n <= 0) return hipSuccess;
    const long long slots = (long long)A.nrows * A.nslots;
    int gs = (int)((slots + 255) / 256);
    if (gs > 4096) gs = 4096;  // grid-stride loops; keeps the launch of a skipped build cheap
    const int* skip = F.csr_state;  // null: always rebuild
    if (F.csr_state) {
        hipLaunchKernelGGL(csr_key_kernel, dim3(1), dim3(64), 0, stream, F.csr_state, A.nrows, A.nslots, F.n_f);
        hipLaunchKernelGGL(csr_compare_kernel, dim3(gs), dim3(256), 0, stream, A.js, F.js_copy, slots, F.csr_state);
    }
    hipLaunchKernelGGL(zero_ints_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, stream, F.cnt, 2 * n, skip);
    hipLaunchKernelGGL(rev_count_kernel, dim3(gs), dim3(256), 0, stream, A.js, A.nrows, A.nslots, F.n_f, F.cnt, skip);
    {
        const int ntiles = (n + kScanTile - 1) / kScanTile;
        int* tile_sum = F.cnt + 2 * n;
        int* tile_off = tile_sum + ntiles;
        hipLaunchKernelGGL(scan_tiles_kernel, dim3(ntiles), dim3(256), 0, stream, F.cnt, n, tile_sum, skip);
        hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(256), 0, stream, tile_sum, ntiles, n, tile_off, F.off, skip);
        hipLaunchKernelGGL(scan_apply_kernel, dim3(ntiles), dim3(256), 0, stream, F.cnt, n, tile_off, F.off, skip);
    }
    hipLaunchKernelGGL(rev_fill_kernel, dim3(gs), dim3(256), 0, stream, A.js, A.nrows, A.nslots, F.n_f, F.off,
                       F.cnt + n, F.rev, skip);
    hipLaunchKernelGGL(force_kernel, dim3((n + 3) / 4), dim3(256), 0, stream, A.g, A.nrows, A.nslots, F.n_f,
                       F.off, F.rev, F.F);
    return hipGetLastError();
}

hipError_t launch_totals(const AtomArgs& A, const ForceArgs& F, hipStream_t stream)
{
    hipLaunchKernelGGL(totals_partial_kernel, dim3(kTotalsBlocks), dim3(256), 0, stream, A.Ei, A.Wi,
                       A.nrows > 0 ? A.nrows : 0, A.nrows > 0 ? F.n_e : 0, F.tot);
    hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(64), 0, stream, F.tot, F.cell_rank, F.volume, F.out, F.status);
    return hipGetLastError();
}

hipError_t launch_forces(const DevPlan& P, const AtomArgs& A, const ForceArgs& F, size_t lds_bytes,
                         bool radial_fp32, hipStream_t stream, hipEvent_t ev_begin, hipEvent_t ev_end, int variant)
{
    if (A.nrows <= 0) return launch_totals(A, F, stream);
    if (ev_begin) hipEventRecord(ev_begin, stream);
    hipError_t e = launch_atoms(P, A, lds_bytes, radial_fp32, stream, variant, A.nrows);
    if (e != hipSuccess) return e;
    if (ev_end) hipEventRecord(ev_end, stream);
    e = launch_gather(A, F, stream);
    if (e != hipSuccess) return e;
    return launch_totals(A, F, stream);
}

}  // namespace mtp
