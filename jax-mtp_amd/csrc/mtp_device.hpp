// Device helpers shared by the evaluation and training kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "mtp_plan.hpp"

namespace mtp {

// The value of v on lane l (a wave-uniform constant), through two readlanes.
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v of the DPP partner lane (dpp_ctrl CTRL, all lanes active), both halves.
template <int CTRL>
__device__ __forceinline__ double dpp_partner(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes of a wavefront (every lane active); every lane gets
// the same value.  Within each row of 16 lanes, DPP moves exchange the sums
// of lanes 1 and 2 apart (quad permutes), then of the two quads of each
// half-row (half-row mirror) and of the two half-rows (row mirror): VALU
// moves, no LDS round trips.  Partners add the same two values, so every
// lane of a row holds the same bits; the four row totals are added in row
// order through readlanes.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_partner<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_partner<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_partner<0x141>(v);  // row_half_mirror
    v += dpp_partner<0x140>(v);  // row_mirror
    return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}

// 1/nu for a gradient block of tensor order nu (1 <= nu <= kMaxNu): the
// hardware reciprocal refined by two Newton steps, r <- r + r (1 - nu r)
// (each squares the relative error of the estimate), in five instructions
// instead of the FP64 division sequence per block and slot.  The result is
// within an ulp of 1/nu.
__device__ __forceinline__ double recip_order(int nu)
{
    const double x = (double)nu;
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// i / d for the (slot, mu) loops, 0 <= i < 2^17 and 1 <= d <= kMaxR, as a
// multiply and shift: with magic = floor(2^20 / d) + 1 = 2^20 / d + e
// (0 < e <= 1), i * magic / 2^20 = i / d + i e / 2^20 and i e / 2^20 < 1/d,
// so the floor is exact there (checked exhaustively for these d and i).
// The divisor is a kernel argument, so a plain division would be the
// generic integer division sequence per item.
__host__ __device__ __forceinline__ unsigned row_quotient_magic(int d) { return (1u << 20) / (unsigned)d + 1u; }
__host__ __device__ __forceinline__ int row_quotient(int i, unsigned magic) { return (int)(((unsigned)i * magic) >> 20); }

__device__ __forceinline__ int wrap_species(int t, int Sp)
{
    if (t < 0) t += Sp;
    if (t < 0) t = 0;
    if (t >= Sp) t = Sp - 1;
    return t;
}

// Radial basis of one slot (jax_backend.py:83-95, 108-160).  Writes f[mu]
// and df[mu] = d f_mu / d r for mu < R.  Real is double or float.
template <typename Real>
__device__ __forceinline__ void radial_slot(const DevPlan& P, double r, int ti, int tk, double* f, double* df)
{
    const Real rmin = (Real)P.min_dist, rmax = (Real)P.max_dist;
    const Real rr = (Real)r;
    const Real range_inv = (Real)1 / (rmax - rmin);
    const Real x = ((Real)2 * rr - (rmin + rmax)) * range_inv;
    const Real dxdr = (Real)2 * range_inv;
    const Real s = (rmax - rr) * (rmax - rr);
    const Real ds = (Real)-2 * (rmax - rr);
    const double* c = P.radial_coeffs + (size_t)(wrap_species(ti, P.Sp) * P.Sp + wrap_species(tk, P.Sp)) * P.R * P.N;
    Real af[kMaxR], ad[kMaxR];
#pragma unroll
    for (int m = 0; m < kMaxR; ++m) {
        af[m] = m < P.R ? (Real)c[m * P.N] : (Real)0;
        ad[m] = 0;
    }
    if (P.N > 1) {
#pragma unroll
        for (int m = 0; m < kMaxR; ++m)
            if (m < P.R) { af[m] += (Real)c[m * P.N + 1] * x; ad[m] += (Real)c[m * P.N + 1]; }
        Real t2 = 1, t1 = x, d2 = 0, d1 = 1;
        for (int n = 2; n < P.N; ++n) {
            const Real t0 = (Real)2 * x * t1 - t2;
            const Real d0 = (Real)2 * t1 + (Real)2 * x * d1 - d2;
#pragma unroll
            for (int m = 0; m < kMaxR; ++m)
                if (m < P.R) { af[m] += (Real)c[m * P.N + n] * t0; ad[m] += (Real)c[m * P.N + n] * d0; }
            t2 = t1; t1 = t0; d2 = d1; d1 = d0;
        }
    }
    const Real sc = (Real)P.scaling;
#pragma unroll
    for (int m = 0; m < kMaxR; ++m) {
        if (m < P.R) {
            f[m] = (double)(sc * s * af[m]);
            df[m] = (double)(sc * (ds * af[m] + s * ad[m] * dxdr));
        }
    }
}

// One radial function of one slot from its N Chebyshev coefficients c:
// f_mu(r) and d f_mu / d r, with the same operations as radial_slot for that
// mu (so the results are identical).  The coefficient loads of a block of eight Chebyshev terms are independent and
// issue together; one thread per (slot, mu) spreads the basis over a
// workgroup.
template <typename Real>
__device__ __forceinline__ void radial_one_c(const DevPlan& P, double r, const double* c, double& f, double& df)
{
    const Real rmin = (Real)P.min_dist, rmax = (Real)P.max_dist;
    const Real rr = (Real)r;
    const Real range_inv = (Real)1 / (rmax - rmin);
    const Real x = ((Real)2 * rr - (rmin + rmax)) * range_inv;
    const Real dxdr = (Real)2 * range_inv;
    const Real s = (rmax - rr) * (rmax - rr);
    const Real ds = (Real)-2 * (rmax - rr);
    Real af = (Real)c[0], ad = 0;
    if (P.N > 1) {
        af += (Real)c[1] * x;
        ad += (Real)c[1];
        Real t2 = 1, t1 = x, d2 = 0, d1 = 1;
#pragma unroll 8
        for (int n = 2; n < P.N; ++n) {
            const Real t0 = (Real)2 * x * t1 - t2;
            const Real d0 = (Real)2 * t1 + (Real)2 * x * d1 - d2;
            af += (Real)c[n] * t0;
            ad += (Real)c[n] * d0;
            t2 = t1; t1 = t0; d2 = d1; d1 = d0;
        }
    }
    const Real sc = (Real)P.scaling;
    f = (double)(sc * s * af);
    df = (double)(sc * (ds * af + s * ad * dxdr));
}

// The same with the coefficients of (ti, tk, mu) read from the plan's global table.
template <typename Real>
__device__ __forceinline__ void radial_one(const DevPlan& P, double r, int ti, int tk, int mu, double& f, double& df)
{
    radial_one_c<Real>(P, r,
                       P.radial_coeffs +
                           ((size_t)(wrap_species(ti, P.Sp) * P.Sp + wrap_species(tk, P.Sp)) * P.R + mu) * P.N,
                       f, df);
}

// FP64 radial basis of one slot with the per-coefficient factors used by the
// parameter gradient: a[n*stride] = d f_mu / d c_n = scaling * s * T_n and
// b[n*stride] = d f'_mu / d c_n = scaling * (s' T_n + s T_n' dx/dr).
__device__ __forceinline__ void radial_slot_coef(const DevPlan& P, double r, int ti, int tk, double* f, double* df,
                                                 double* a, double* b, int stride)
{
    const double rmin = P.min_dist, rmax = P.max_dist;
    const double range_inv = 1.0 / (rmax - rmin);
    const double x = (2.0 * r - (rmin + rmax)) * range_inv;
    const double dxdr = 2.0 * range_inv;
    const double s = (rmax - r) * (rmax - r);
    const double ds = -2.0 * (rmax - r);
    const double sc = P.scaling;
    const double* c = P.radial_coeffs + (size_t)(wrap_species(ti, P.Sp) * P.Sp + wrap_species(tk, P.Sp)) * P.R * P.N;
    double af[kMaxR], ad[kMaxR];
#pragma unroll
    for (int m = 0; m < kMaxR; ++m) { af[m] = 0.0; ad[m] = 0.0; }
    double t2 = 0.0, t1 = 0.0, d2 = 0.0, d1 = 0.0;
    for (int n = 0; n < P.N; ++n) {
        double t0, d0;
        if (n == 0) { t0 = 1.0; d0 = 0.0; }
        else if (n == 1) { t0 = x; d0 = 1.0; }
        else { t0 = 2.0 * x * t1 - t2; d0 = 2.0 * t1 + 2.0 * x * d1 - d2; }
        a[n * stride] = sc * s * t0;
        b[n * stride] = sc * (ds * t0 + s * d0 * dxdr);
#pragma unroll
        for (int m = 0; m < kMaxR; ++m)
            if (m < P.R) { af[m] += c[m * P.N + n] * t0; ad[m] += c[m * P.N + n] * d0; }
        t2 = t1; t1 = t0; d2 = d1; d1 = d0;
    }
#pragma unroll
    for (int m = 0; m < kMaxR; ++m)
        if (m < P.R) { f[m] = sc * s * af[m]; df[m] = sc * (ds * af[m] + s * ad[m] * dxdr); }
}

// Canonical index of monomial (a, b, NU-a-b) inside a degree-NU block.
__host__ __device__ constexpr int mono_idx(int NU, int a, int b) { return (NU - a) * (NU - a + 1) / 2 + (NU - a - b); }

// H = sum_alpha d[alpha] grad_u u^alpha over one (mu, NU) block, written as a
// sum over the degree NU-1 monomials (Euler: u . H = NU * sum d P).
template <int NU, int DS = 1>
__device__ __forceinline__ void grad_block(const double* __restrict__ d, const double* px, const double* py,
                                           const double* pz, double& Hx, double& Hy, double& Hz)
{
#pragma unroll
    for (int a = NU - 1; a >= 0; --a) {
#pragma unroll
        for (int b = NU - 1 - a; b >= 0; --b) {
            const int c = NU - 1 - a - b;
            const double P = px[a] * py[b] * pz[c];
            Hx += (double)(a + 1) * d[DS * mono_idx(NU, a + 1, b)] * P;
            Hy += (double)(b + 1) * d[DS * mono_idx(NU, a, b + 1)] * P;
            Hz += (double)(c + 1) * d[DS * mono_idx(NU, a, b)] * P;
        }
        // one row of monomials at a time: keeps the LDS loads of the fully
        // unrolled block from being hoisted together (register pressure)
        __builtin_amdgcn_sched_barrier(0);
    }
}

// grad_block without a z power table: along a row of monomials (fixed a,
// b decreasing) c increases by one, so u_z^c is carried as a running
// product, built by the same multiplications as the table entries (the
// results are identical).  Frees PN registers of doubles.
template <int NU, int DS = 1>
__device__ __forceinline__ void grad_block_xy(const double* __restrict__ d, const double* px, const double* py,
                                              double uz, double& Hx, double& Hy, double& Hz)
{
#pragma unroll
    for (int a = NU - 1; a >= 0; --a) {
        double pzc = 1.0;
#pragma unroll
        for (int b = NU - 1 - a; b >= 0; --b) {
            const int c = NU - 1 - a - b;
            const double P = px[a] * py[b] * pzc;
            Hx += (double)(a + 1) * d[DS * mono_idx(NU, a + 1, b)] * P;
            Hy += (double)(b + 1) * d[DS * mono_idx(NU, a, b + 1)] * P;
            Hz += (double)(c + 1) * d[DS * mono_idx(NU, a, b)] * P;
            pzc *= uz;
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int PN>
__device__ __forceinline__ void opaque_powers_xy(double* px, double* py)
{
#pragma unroll
    for (int p = 0; p < PN; ++p) asm volatile("" : "+v"(px[p]), "+v"(py[p]));
}

template <int PN, int DS = 1>
__device__ __forceinline__ void grad_block_dispatch_xy(int nu, const double* d, double* px, double* py, double uz,
                                                       double& Hx, double& Hy, double& Hz)
{
    opaque_powers_xy<PN>(px, py);
    switch (nu) {
        case 1: grad_block_xy<1, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 2: if constexpr (PN >= 2) grad_block_xy<2, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 3: if constexpr (PN >= 3) grad_block_xy<3, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 4: if constexpr (PN >= 4) grad_block_xy<4, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 5: if constexpr (PN >= 5) grad_block_xy<5, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 6: if constexpr (PN >= 6) grad_block_xy<6, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 7: if constexpr (PN >= 7) grad_block_xy<7, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 8: if constexpr (PN >= 8) grad_block_xy<8, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 9: if constexpr (PN >= 9) grad_block_xy<9, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 10: if constexpr (PN >= 10) grad_block_xy<10, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 11: if constexpr (PN >= 11) grad_block_xy<11, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        case 12: if constexpr (PN >= 12) grad_block_xy<12, DS>(d, px, py, uz, Hx, Hy, Hz); break;
        default: break;
    }
}

// Called once per (mu, nu) block inside a loop over blocks.  The monomial
// products only depend on the slot, so the compiler would hoist all of them
// (every order 1..12) out of that loop and spill; the empty asm marks the
// power tables as rewritten each time, which keeps the products per block.
// PN is the length of the power tables (orders nu <= PN are handled).
template <int PN = kMaxNu + 1>
__device__ __forceinline__ void opaque_powers(double* px, double* py, double* pz)
{
#pragma unroll
    for (int p = 0; p < PN; ++p) asm volatile("" : "+v"(px[p]), "+v"(py[p]), "+v"(pz[p]));
}

// DS is the stride of the adjoints in d (2 when two rows are interleaved).
template <int PN = kMaxNu + 1, int DS = 1>
__device__ __forceinline__ void grad_block_dispatch(int nu, const double* d, double* px, double* py, double* pz,
                                                    double& Hx, double& Hy, double& Hz)
{
    opaque_powers<PN>(px, py, pz);
    switch (nu) {
        case 1: grad_block<1, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 2: if constexpr (PN >= 2) grad_block<2, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 3: if constexpr (PN >= 3) grad_block<3, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 4: if constexpr (PN >= 4) grad_block<4, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 5: if constexpr (PN >= 5) grad_block<5, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 6: if constexpr (PN >= 6) grad_block<6, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 7: if constexpr (PN >= 7) grad_block<7, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 8: if constexpr (PN >= 8) grad_block<8, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 9: if constexpr (PN >= 9) grad_block<9, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 10: if constexpr (PN >= 10) grad_block<10, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 11: if constexpr (PN >= 11) grad_block<11, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        case 12: if constexpr (PN >= 12) grad_block<12, DS>(d, px, py, pz, Hx, Hy, Hz); break;
        default: break;
    }
}

}  // namespace mtp
