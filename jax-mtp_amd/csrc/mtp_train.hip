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

// ---------------------------------------------------------------------------
// Dual contraction streams.  Each step of the atom kernels' streams holds up
// to three entries of a lane with one integer multiplier w (operand slots
// a, b per entry; the last step of the lane's item flagged in the
// multiplier word, which then also holds the destination slot), and every
// entry reads two 16-byte pairs.  Forward: a = (m_a, mt_a), b = (m_b, mt_b) and the sum
// gains w (m_a m_b, mt_a m_b + m_a mt_b).  Reverse: a = (m_o, mt_o) of the
// operand, b = (dm_t, dmt_t) of the target it feeds, and the sum gains
// w (m_o dm_t + mt_o dmt_t, m_o dmt_t).
// ---------------------------------------------------------------------------
template <bool REV>
__device__ __forceinline__ double2 dual_term(const double2 a, const double2 b, double2 t)
{
    if (REV) {
        t.x = fma(a.x, b.x, fma(a.y, b.y, t.x));
        t.y = fma(a.x, b.y, t.y);
    } else {
        t.x = fma(a.x, b.x, t.x);
        t.y = fma(a.x, b.y, fma(a.y, b.x, t.y));
    }
    return t;
}

template <bool REV>
__device__ __forceinline__ void dual_steps(const uint4* q, char* lb, double2& acc)
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
        double2 t = dual_term<REV>(a0[u], b0[u], make_double2(0.0, 0.0));
        t = dual_term<REV>(a1[u], b1[u], t);
        t = dual_term<REV>(a2[u], b2[u], t);
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

// One dual contraction layer (forward: m2 targets; reverse: dm2 targets,
// seeded with (alpha xi, xi) for the mapped moments).
template <int NT, bool REV>
__device__ __forceinline__ void dual_layer(const DevPlan& P, int L, double* lds, double2* m2, double2* dm2,
                                           const double2* part, int tid, double alpha)
{
    constexpr int NW = NT / kWave;
    constexpr int U = kStreamUnroll;
    const int lane = tid & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
    const I4 wv = (REV ? P.rev_wave : P.fwd_wave)[L * NW + w];
    const int nsteps = wv.y;
    // this thread's first combination, base-only and xi-added rows: their
    // table entries and xi values (dependent global loads) are issued before
    // the stream walk, which hides their latency
    const int c0 = REV ? P.rev_cb_beg[L] : P.fwd_cb_beg[L];
    const int c1 = REV ? P.rev_cb_beg[L + 1] : P.fwd_cb_beg[L + 1];
    const I4* comb = REV ? P.rev_comb : P.fwd_comb;
    const int cA = c0 + tid;
    I4 cbA = {0, 0, 0, 0};
    double xA = 0.0;
    if (cA < c1) {
        cbA = comb[cA];
        if (REV) xA = P.xi_of[cbA.x];
    }
    int boX = -1, xaX = -1;
    double boV = 0.0, xaV = 0.0;
    if (REV) {
        const int tb0 = P.rev_bo_beg[L] + tid, ta0 = P.rev_xa_beg[L] + tid;
        if (tb0 < P.rev_bo_beg[L + 1]) { boX = P.rev_bo[tb0]; boV = P.xi_of[boX]; }
        if (ta0 < P.rev_xa_beg[L + 1]) { xaX = P.rev_xa[ta0]; xaV = P.xi_of[xaX]; }
    }
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
            dual_steps<REV>(qa, lb, acc);
            s0 += U;
            if (s0 >= nsteps) break;
            if (s0 + U < nsteps) {
#pragma unroll
                for (int u = 0; u < U; ++u) qa[u] = tb[(size_t)(s0 + U + u) * kWave];
            }
            dual_steps<REV>(qb, lb, acc);
            s0 += U;
            if (s0 >= nsteps) break;
        }
    }
    __syncthreads();
    double2* dst = REV ? dm2 : m2;
    for (int c = cA; c < c1; c += NT) {
        const I4 cb = c == cA ? cbA : comb[c];
        const double x0 = REV ? (c == cA ? xA : P.xi_of[cb.x]) : 0.0;
        double2 acc = make_double2(alpha * x0, x0);
        for (int p = cb.y; p < cb.z; ++p) {
            const double2 v = part[p];
            acc.x += v.x;
            acc.y += v.y;
        }
        dst[cb.x] = acc;
    }
    if (REV) {
        for (int t = P.rev_bo_beg[L] + tid; t < P.rev_bo_beg[L + 1]; t += NT) {
            const bool first = t == P.rev_bo_beg[L] + tid;
            const int x = first ? boX : P.rev_bo[t];
            const double xi = first ? boV : P.xi_of[x];
            dm2[x] = make_double2(alpha * xi, xi);
        }
        for (int t = P.rev_xa_beg[L] + tid; t < P.rev_xa_beg[L + 1]; t += NT) {
            const bool first = t == P.rev_xa_beg[L] + tid;
            const int x = first ? xaX : P.rev_xa[t];
            const double xi = first ? xaV : P.xi_of[x];
            double2 v = dm2[x];
            v.x += alpha * xi;
            v.y += xi;
            dm2[x] = v;
        }
    }
    if (c1 > c0 || REV) __syncthreads();
}

// Basic moment of monomial (a, b, c) and its tangent over slots [0, n), for
// mu < NMU: acc[mu] = sum_k f_mu(k) P_k, acct[mu] = sum_k fu_mu(k) P_k +
// f_mu(k) G_k with P = u_x^a u_y^b u_z^c and G = grad_u P . e_k (tables
// px.. of the powers, ex.. of p u^(p-1) e; f and fu at stride R).  Slots go
// in batches whose LDS reads issue before their arithmetic.
template <int NMU>
__device__ __forceinline__ void dual_moment_sum(const double* px, const double* py, const double* pz,
                                                const double* ex, const double* ey, const double* ez,
                                                const double* f, const double* fu, int R, int n, double* acc,
                                                double* acct)
{
    constexpr int UB = NMU <= 2 ? 4 : 2;
    int k = 0;
    for (; k + UB <= n; k += UB) {
        double a[UB], b[UB], c[UB], da[UB], db[UB], dc[UB], fr[UB][NMU], fur[UB][NMU];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            a[u] = px[k + u]; b[u] = py[k + u]; c[u] = pz[k + u];
            da[u] = ex[k + u]; db[u] = ey[k + u]; dc[u] = ez[k + u];
#pragma unroll
            for (int mu = 0; mu < NMU; ++mu) { fr[u][mu] = f[(k + u) * R + mu]; fur[u][mu] = fu[(k + u) * R + mu]; }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const double bc = b[u] * c[u];
            const double Pm = a[u] * bc;
            const double G = fma(da[u], bc, a[u] * fma(db[u], c[u], b[u] * dc[u]));
#pragma unroll
            for (int mu = 0; mu < NMU; ++mu) {
                acc[mu] = fma(fr[u][mu], Pm, acc[mu]);
                acct[mu] = fma(fur[u][mu], Pm, fma(fr[u][mu], G, acct[mu]));
            }
        }
    }
    for (; k < n; ++k) {
        const double bc = py[k] * pz[k];
        const double Pm = px[k] * bc;
        const double G = fma(ex[k], bc, px[k] * fma(ey[k], pz[k], py[k] * ez[k]));
#pragma unroll
        for (int mu = 0; mu < NMU; ++mu) {
            acc[mu] = fma(f[k * R + mu], Pm, acc[mu]);
            acct[mu] = fma(fu[k * R + mu], Pm, fma(f[k * R + mu], G, acct[mu]));
        }
    }
}

// H = sum_alpha d_alpha grad_u u^alpha and T = the same over dt, for one
// (mu, NU) block of (dm, dmt) pairs (u_z^c carried as a running product).
template <int NU>
__device__ __forceinline__ void dual_grad_block(const double2* __restrict__ d, const double* px, const double* py,
                                                double uz, double* H, double* T)
{
#pragma unroll
    for (int a = NU - 1; a >= 0; --a) {
        double pzc = 1.0;
#pragma unroll
        for (int b = NU - 1 - a; b >= 0; --b) {
            const int c = NU - 1 - a - b;
            const double P = px[a] * py[b] * pzc;
            const double2 dx = d[mono_idx(NU, a + 1, b)];
            const double2 dy = d[mono_idx(NU, a, b + 1)];
            const double2 dz = d[mono_idx(NU, a, b)];
            const double Pa = (double)(a + 1) * P, Pb = (double)(b + 1) * P, Pc = (double)(c + 1) * P;
            H[0] = fma(dx.x, Pa, H[0]); T[0] = fma(dx.y, Pa, T[0]);
            H[1] = fma(dy.x, Pb, H[1]); T[1] = fma(dy.y, Pb, T[1]);
            H[2] = fma(dz.x, Pc, H[2]); T[2] = fma(dz.y, Pc, T[2]);
            pzc *= uz;
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int PN>
__device__ __forceinline__ void dual_grad_dispatch(int nu, const double2* d, double* px, double* py, double uz,
                                                   double* H, double* T)
{
    opaque_powers_xy<PN>(px, py);
    switch (nu) {
        case 1: dual_grad_block<1>(d, px, py, uz, H, T); break;
        case 2: if constexpr (PN >= 2) dual_grad_block<2>(d, px, py, uz, H, T); break;
        case 3: if constexpr (PN >= 3) dual_grad_block<3>(d, px, py, uz, H, T); break;
        case 4: if constexpr (PN >= 4) dual_grad_block<4>(d, px, py, uz, H, T); break;
        case 5: if constexpr (PN >= 5) dual_grad_block<5>(d, px, py, uz, H, T); break;
        case 6: if constexpr (PN >= 6) dual_grad_block<6>(d, px, py, uz, H, T); break;
        case 7: if constexpr (PN >= 7) dual_grad_block<7>(d, px, py, uz, H, T); break;
        case 8: if constexpr (PN >= 8) dual_grad_block<8>(d, px, py, uz, H, T); break;
        case 9: if constexpr (PN >= 9) dual_grad_block<9>(d, px, py, uz, H, T); break;
        case 10: if constexpr (PN >= 10) dual_grad_block<10>(d, px, py, uz, H, T); break;
        case 11: if constexpr (PN >= 11) dual_grad_block<11>(d, px, py, uz, H, T); break;
        case 12: if constexpr (PN >= 12) dual_grad_block<12>(d, px, py, uz, H, T); break;
        default: break;
    }
}

// Slot k, part prt: over the part's (mu, nu) blocks, A = sum dm P
// (d J / d f_mu), Ap = sum dmt P (d J / d f'_mu, times u . Delta) and
// Q = sum dmt grad_r P . Delta (d J / d f_mu through the tangent), with
// grad_r P . Delta = T . e for e = (Delta - (u . Delta) u) / r.  Adds
// A + Q to acc[mu] and Ap to acc[R + mu] (the factor u . Delta is in cb).
template <int PN>
__device__ __forceinline__ void dual_slot(const DevPlan& P, int prt, double ux, double uy, double uz, double e0,
                                          double e1, double e2, const double2* dm2, double* acc)
{
    double px[PN], py[PN];
    px[0] = py[0] = 1.0;
#pragma unroll
    for (int p = 1; p < PN; ++p) { px[p] = px[p - 1] * ux; py[p] = py[p - 1] * uy; }
    const int R = P.R;
    const int b1 = P.part_beg[prt + 1];
    for (int bi = P.part_beg[prt]; bi < b1; ++bi) {
        const I2 bl = P.part_blk[bi];
        const int mu = bl.x & 0xff, nu = bl.y;
        const double2* d = dm2 + (bl.x >> 8);
        double a0, a1, q;
        if (nu == 0) {
            const double2 v = d[0];
            a0 = v.x; a1 = v.y; q = 0.0;
        } else {
            double H[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
            dual_grad_dispatch<PN>(nu, d, px, py, uz, H, T);
            const double inu = recip_order(nu);
            a0 = (ux * H[0] + uy * H[1] + uz * H[2]) * inu;
            a1 = (ux * T[0] + uy * T[1] + uz * T[2]) * inu;
            q = T[0] * e0 + T[1] * e1 + T[2] * e2;
        }
        acc[mu] += a0 + q;
        acc[R + mu] += a1;
    }
}

// One row (atom) per workgroup, one thread per slot in the compaction.
static_assert(kTrainThreads >= kAtomBlock, "rows hold at most kAtomBlock slots");
__global__ void __launch_bounds__(kTrainThreads) __attribute__((amdgpu_waves_per_eu(3)))
train_kernel(const DevPlan P, TrainArgs A)
{
    constexpr int NT = kTrainThreads;
    constexpr int NW = NT / kWave;
    extern __shared__ double lds[];
    const int row = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int K = A.nslots, KL = P.k_layout, R = P.R, N = P.N;
    const TrainLds L = train_lds_layout(P.amc, P.n_live, P.B, KL, R, N, P.nu_max, P.n_parts, NT, P.max_partials);
    double2* l2 = reinterpret_cast<double2*>(lds);
    double2* m2 = l2 + L.m;
    double2* dm2 = l2 + L.dm;
    const double2* part2 = l2 + L.part;
    double* vx = lds + L.vx;
    double* vy = lds + L.vy;
    double* vz = lds + L.vz;
    double* rr = lds + L.r;
    double* ex = lds + L.ex;
    double* ey = lds + L.ey;
    double* ez = lds + L.ez;
    double* ud = lds + L.ud;
    double* ca = lds + L.ca;
    double* cb = lds + L.cb;
    int* ints = reinterpret_cast<int*>(lds + L.ints);
    int* jtc = ints;              // [KL] neighbor type of compact slot
    int* wave_cnt = ints + KL;    // [NW]
    int* nvp = wave_cnt + NW;
    if (tid == 0) l2[0] = make_double2(0.0, 0.0);

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

    // ---- A: compaction of the valid slots; per slot the perturbation
    // Delta = gF_i - gF_j + Lambda r, u . Delta and e = (Delta - (u . Delta) u) / r
    int nv;
    {
        const int k = tid;
        double v0 = 0, v1 = 0, v2 = 0, r = 1.0;
        double dx = gFx, dy = gFy, dz = gFz;
        int jt = -1;
        bool valid = false;
        if (k < K) {
            // every load of the slot (the neighbor's gF included) issues
            // before the compaction's barrier
            v0 = rij_row[3 * k]; v1 = rij_row[3 * k + 1]; v2 = rij_row[3 * k + 2];
            jt = jt_row[k];
            const int j = js_row[k];
            if (j >= 0 && j < A.nrows) {
                dx -= A.gF[3 * (size_t)j]; dy -= A.gF[3 * (size_t)j + 1]; dz -= A.gF[3 * (size_t)j + 2];
            }
            r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
            valid = (jt >= 0) && (r < P.max_dist);
        }
        const uint64_t mask = __ballot(valid);
        const int wave = tid / kWave;
        if (lane == 0) wave_cnt[wave] = __popcll(mask);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
        if (valid) {
            jtc[pos] = jt;
            dx += Lxx * v0 + Lxy * v1 + Lxz * v2;
            dy += Lxy * v0 + Lyy * v1 + Lyz * v2;
            dz += Lxz * v0 + Lyz * v1 + Lzz * v2;
            const double ir = 1.0 / r;
            const double ux = v0 * ir, uy = v1 * ir, uz = v2 * ir;
            const double udv = ux * dx + uy * dy + uz * dz;
            vx[pos] = v0; vy[pos] = v1; vz[pos] = v2; rr[pos] = r;
            ud[pos] = udv;
            ex[pos] = (dx - udv * ux) * ir;
            ey[pos] = (dy - udv * uy) * ir;
            ez[pos] = (dz - udv * uz) * ir;
        }
        if (tid == 0) {
            int tot = 0;
            for (int w = 0; w < NW; ++w) tot += wave_cnt[w];
            *nvp = tot;
        }
        __syncthreads();
        nv = __builtin_amdgcn_readfirstlane(*nvp);
    }

    // radial values f and fu = f' (u . Delta) per (slot, mu); coefficient
    // factors ca[n][k] = d f / d c_n, cb[n][k] = (d f' / d c_n) (u . Delta)
    // per (slot, n); power tables pw[(q np1 + p) Kp + k] = u_q^p and
    // pe[..] = p u_q^(p-1) e_q per (slot, component)
    const int np1 = P.nu_max + 1;
    const int Kp = KL | 1;
    double* pw = lds + L.pw;
    double* pe = lds + L.pe;
    double* f = lds + L.f;
    double* fu = lds + L.fu;
    for (int i = tid; i < nv * R; i += NT) {
        const int c = i / R, mu = i - c * R;
        double fv, dfv;
        radial_one<double>(P, rr[c], ti, jtc[c], mu, fv, dfv);
        f[i] = fv;
        fu[i] = dfv * ud[c];
    }
    {
        const double rmin = P.min_dist, rmax = P.max_dist;
        const double range_inv = 1.0 / (rmax - rmin);
        const double dxdr = 2.0 * range_inv;
        const double sc = P.scaling;
        for (int i = tid; i < nv * N; i += NT) {
            const int c = i / N, n = i - c * N;
            const double r = rr[c];
            const double x = (2.0 * r - (rmin + rmax)) * range_inv;
            const double s = (rmax - r) * (rmax - r);
            const double ds = -2.0 * (rmax - r);
            double t0 = 1.0, d0 = 0.0;
            if (n >= 1) {
                double t2 = 1.0, t1 = x, d2 = 0.0, d1 = 1.0;
                t0 = x; d0 = 1.0;
                for (int q = 2; q <= n; ++q) {
                    t0 = 2.0 * x * t1 - t2;
                    d0 = 2.0 * t1 + 2.0 * x * d1 - d2;
                    t2 = t1; t1 = t0; d2 = d1; d1 = d0;
                }
            }
            ca[n * KL + c] = sc * s * t0;
            cb[n * KL + c] = sc * (ds * t0 + s * d0 * dxdr) * ud[c];
        }
    }
    for (int t = tid; t < nv * 3; t += NT) {
        const int k = t / 3, q = t - 3 * k;
        const double u = (q == 0 ? vx[k] : (q == 1 ? vy[k] : vz[k])) / rr[k];
        const double e = q == 0 ? ex[k] : (q == 1 ? ey[k] : ez[k]);
        double* col = pw + (size_t)q * np1 * Kp + k;
        double* dcol = pe + (size_t)q * np1 * Kp + k;
        double v = 1.0;
        col[0] = 1.0;
        dcol[0] = 0.0;
        for (int p = 1; p < np1; ++p) {
            dcol[(size_t)p * Kp] = (double)p * v * e;
            v *= u;
            col[(size_t)p * Kp] = v;
        }
    }
    __syncthreads();

    // ---- B: basic moments and their tangents, one thread per monomial
    // (all its mu; the mu range is the largest of the monomial's wavefront)
    {
        const int ngrp = (P.n_mono + kWave - 1) / kWave;
        for (int g = tid / kWave; g < ngrp; g += NW) {
            const int t = g * kWave + lane;
            if (t < P.n_mono) {
                const I4 mo = P.mono[t];
                const int gmax = __builtin_amdgcn_readfirstlane(mo.w >> 24);
                const double* px = pw + (size_t)mo.x * Kp;
                const double* py = pw + (size_t)(np1 + mo.y) * Kp;
                const double* pz = pw + (size_t)(2 * np1 + mo.z) * Kp;
                const double* qx = pe + (size_t)mo.x * Kp;
                const double* qy = pe + (size_t)(np1 + mo.y) * Kp;
                const double* qz = pe + (size_t)(2 * np1 + mo.z) * Kp;
                double acc[kMaxR], acct[kMaxR];
#pragma unroll
                for (int mu = 0; mu < kMaxR; ++mu) { acc[mu] = 0.0; acct[mu] = 0.0; }
                switch (gmax) {
                    case 0: dual_moment_sum<1>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    case 1: dual_moment_sum<2>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    case 2: dual_moment_sum<3>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    case 3: dual_moment_sum<4>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    case 4: dual_moment_sum<5>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    case 5: dual_moment_sum<6>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    case 6: dual_moment_sum<7>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                    default: dual_moment_sum<8>(px, py, pz, qx, qy, qz, f, fu, R, nv, acc, acct); break;
                }
                const I4 d0 = P.mono_dst[2 * t], d1 = P.mono_dst[2 * t + 1];
                const int dst[kMaxR] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
                for (int mu = 0; mu < kMaxR; ++mu)
                    if (dst[mu] >= 0) m2[dst[mu]] = make_double2(acc[mu], acct[mu]);
            }
        }
    }
    __syncthreads();
    // moments that are never targets stay 0 (m2[x >= n_live] lies in the
    // adjoint region, free once the tables are consumed)
    for (int x = P.B + tid; x < P.amc; x += NT) m2[x] = make_double2(0.0, 0.0);
    __syncthreads();

    // ---- C: contraction layers for (m, mt)
    for (int d = 0; d < P.n_depth; ++d) dual_layer<NT, false>(P, d, lds, m2, dm2, part2, tid, alpha);

    // ---- D: dJ/dxi_s of this atom
    for (int s = tid; s < P.S; s += NT) {
        const double2 v = m2[P.mapping[s]];
        A.gxi_part[(size_t)row * P.S + s] = alpha * v.x + v.y;
    }
    __syncthreads();  // never-operand moments alias the adjoints written next

    // ---- E: reverse layers over (m, mt): (dm, dmt) = (dJ/dm, dJ/dmt)
    for (int li = 0; li <= P.n_depth; ++li) dual_layer<NT, true>(P, li, lds, m2, dm2, part2, tid, alpha);

    // ---- F: per (slot, part) accumulators, past the basic moments' adjoints
    double* acc = lds + L.acc;
    for (int t = tid; t < P.n_parts * KL * 2 * R; t += NT) acc[t] = 0.0;
    __syncthreads();
    {
        const int kk = tid & (kWave - 1), prt = tid / kWave;
        for (int k = kk; k < nv && prt < P.n_parts; k += kWave) {
            const double ir = 1.0 / rr[k];
            const double ux = vx[k] * ir, uy = vy[k] * ir, uz = vz[k] * ir;
            double* o = acc + ((size_t)prt * KL + k) * 2 * R;
            const double e0 = ex[k], e1 = ey[k], e2 = ez[k];
            const int nu_max = P.nu_max;
            if (nu_max <= 4) dual_slot<4>(P, prt, ux, uy, uz, e0, e1, e2, dm2, o);
            else if (nu_max <= 6) dual_slot<6>(P, prt, ux, uy, uz, e0, e1, e2, dm2, o);
            else if (nu_max <= 8) dual_slot<8>(P, prt, ux, uy, uz, e0, e1, e2, dm2, o);
            else if (nu_max <= 10) dual_slot<10>(P, prt, ux, uy, uz, e0, e1, e2, dm2, o);
            else dual_slot<kMaxNu>(P, prt, ux, uy, uz, e0, e1, e2, dm2, o);
        }
    }
    __syncthreads();
    // parts summed in order into part 0
    for (int t = tid; t < nv * 2 * R; t += NT) {
        double s = acc[t];
        for (int prt = 1; prt < P.n_parts; ++prt) s += acc[(size_t)prt * KL * 2 * R + t];
        acc[t] = s;
    }
    __syncthreads();

    // dJ/dc[ti][tk][mu][n] of this atom (column (tk R + mu) N + n), slots
    // split over G thread groups whose sums are added in group order; plus
    // alpha for the species gradient
    {
        const int ncols = P.Sp * R * N;
        double* out = A.gc_part + (size_t)row * (1 + ncols);
        double* red = lds + L.red;
        if (tid == 0) out[0] = alpha;
        const int G = ncols >= NT ? 1 : (NT / ncols < 8 ? NT / ncols : 8);
        for (int c0 = 0; c0 < ncols; c0 += NT / G) {
            const int c = c0 + tid % (NT / G), g = tid / (NT / G);
            double s = 0.0;
            if (c < ncols && g < G) {
                const int n = c % N, mu = (c / N) % R, tk = c / (N * R);
                for (int k = g; k < nv; k += G) {
                    const double v = acc[k * 2 * R + mu] * ca[n * KL + k] + acc[k * 2 * R + R + mu] * cb[n * KL + k];
                    s += wrap_species(jtc[k], P.Sp) == tk ? v : 0.0;
                }
            }
            if (G == 1) {
                if (c < ncols) out[1 + c] = s;
            } else {
                if (g < G) red[g * (NT / G) + tid % (NT / G)] = s;
                __syncthreads();
                if (tid < NT / G && c0 + tid < ncols) {
                    double t = red[tid];
                    for (int q = 1; q < G; ++q) t += red[q * (NT / G) + tid];
                    out[1 + c0 + tid] = t;
                }
                __syncthreads();
            }
        }
    }
}

// Column sums over row chunks; with rowtype the sums are split by type.
// partial[chunk][t][col], chunk = kColsumChunk rows, fixed summation order.
__global__ void __launch_bounds__(256) colsum_stage1(const double* X, int rows, int cols, const int* rowtype, int ntypes,
                                                     int Sp, double* partial)
{
    __shared__ double buf[4][64];
    const int col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int lane = threadIdx.x >> 6;
    const int chunk = blockIdx.y;
    const int r0 = chunk * kColsumChunk, r1 = min(rows, r0 + kColsumChunk);
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

__global__ void __launch_bounds__(256) colsum_stage2(const double* partial, int nchunk, int width, double* out)
{
    // one block per column: strided per-thread sums, then a tree in LDS
    __shared__ double buf[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int q = threadIdx.x; q < nchunk; q += 256) s += partial[(size_t)q * width + c];
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) buf[threadIdx.x] += buf[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = buf[0];
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
    hipLaunchKernelGGL(train_kernel, dim3(A.nrows), dim3(kTrainThreads), lds_bytes, stream, P, A);
    const int nchunk = (A.nrows + kColsumChunk - 1) / kColsumChunk;
    // xi columns (no types)
    hipLaunchKernelGGL(colsum_stage1, dim3((P.S + 63) / 64, nchunk), dim3(256), 0, stream, A.gxi_part, A.nrows, P.S,
                       (const int*)nullptr, 1, P.Sp, scratch_partial);
    hipLaunchKernelGGL(colsum_stage2, dim3(P.S), dim3(256), 0, stream, scratch_partial, nchunk, P.S, xi_sum);
    // alpha + radial columns, split by the row's species
    const int w = 1 + ncols;
    hipLaunchKernelGGL(colsum_stage1, dim3((w + 63) / 64, nchunk), dim3(256), 0, stream, A.gc_part, A.nrows, w,
                       A.itypes, P.Sp, P.Sp, scratch_partial);
    hipLaunchKernelGGL(colsum_stage2, dim3(P.Sp * w), dim3(256), 0, stream, scratch_partial, nchunk, P.Sp * w,
                       typed_sum);
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
