// Launch interface of the HIP kernels (mtp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mtp_plan.hpp"

namespace mtp {

// Padded per-row inputs/outputs of one evaluation (the array contract of
// calc_energy_forces_stress_padded, jax_backend.py:374-421, in FP64).
struct AtomArgs {
    int nrows, nslots;     // A rows, K slots per row
    const double* rij;     // [A][K][3]
    const int* itypes;     // [A]
    const int* jtypes;     // [A][K]
    const int* js;         // [A][K]
    double* Ei;            // [A]      site energies
    double* g;             // [A][K][3] dE_i/dr_ik
    double* Wi;            // [A][9]   per-atom sum_k r_ik (x) g_ik
    // Optional modes of the balanced-item kernels (zero / null for a plain
    // force evaluation):
    const double* xi_atom; // [rows][xi_stride] per-row moment coefficients (file order), replacing xi
    int xi_stride;         // 0: the same vector for every row
    double* basis;         // [A][S] basis values m[mapping], written when not null
    int fwd_only;          // stop after the basis and the energy (g, Wi untouched)
    const int* rows;       // rows to evaluate (null: 0 .. nrows-1)
    const int* row_count;  // device-side length of `rows` (null: nrows)
    // Moment cache of a descriptor pass followed by a per-row-coefficient
    // pass over the same rows (streams variants only): the first writes
    // m[0, n_live) of every row to mom_out [rows][n_live]; the second reads
    // them from mom_in and skips the radial basis, basic moments, forward
    // layers and energy (Ei and basis must be null then; kernels of more
    // than 256 threads recompute instead).
    double* mom_out;
    const double* mom_in;
    const int* skip_row;   // rows with skip_row[row] != 0 are left to another kernel (null: none)
    int* status;           // non-null: set to 1 when a slot vector is not finite (read and cleared by the totals)
};

struct ForceArgs {
    int n_e, n_f, cell_rank;
    double volume;
    int* cnt;              // [2A + 2 ceil(A / 4096) + 16] scratch (counts, cursors, scan tiles)
    int* off;              // [A+1]
    int* rev;              // [A*K]
    double* F;             // [A][3]
    double* out;           // [14]: E, stress6, virial6, status
    double* tot;           // [kTotalsBlocks * 10] scratch of the totals reduction
    // Optional reuse of the reverse-neighbor CSR (cnt, off, rev) of the
    // previous call with the same js, n_f and shape (null: rebuild always).
    int* csr_state;        // [8]: same flag, last nrows, nslots, n_f, valid
    int* js_copy;          // [A][K] js of the last build
    int* status;           // the atom kernels' non-finite-input flag (null: none); the totals clear it
};
constexpr int kTotalsBlocks = 128;

struct AtomLds {
    int m, dm, f, fp, vx, vy, vz, r, red, ints, total_doubles;
    int mlen, dlen;  // doubles of the moment and adjoint regions (in LDS, or a global slice)
};

struct AtomLds2 {
    int m, dm, part, f, gp, gf, gfp, vx, vy, vz, r, red, xl, ints, coef, epart, total_doubles;
};

// LDS of the balanced-item kernels (doubles).  Regions are reused across
// phases:
//   zero [kLdsM] lds[0] = 0.0: operand of the padding entries of the
//               contraction streams and the sink of lanes without an item
//   m   [n_live] moments [0, n_live) (basic or read as operands)
//   dm  [dlen]  A/B: power table (row stride pw_stride(K): slot pairs,
//               an odd number of them, so that a wave's reads of different
//               powers fall in different LDS banks), then f[k * Rp + mu]
//               (Rp = f_stride(R)); C/D: the values of the never-operand
//               moments (m[x] for x >= n_live aliases dm[x - n_live], dead
//               after D); E/F:
//               adjoints, of which F reads only the basic ones [0, B): from
//               max(B, 3K) on, F keeps the per-part gradients and the
//               recomputed f, f'; after F, dm[0, 3K) holds the slot
//               gradients
//   part, vx, vy, vz, r, xl (per-row coefficients), ints, coef (the radial
//               coefficients of the row's species, when staged), epart (the
//               per-wave sums of the site energy, read at the row's end)
// Without tables (a per-row pass over cached moments, which skips phases
// A's radial values and B) the power table and f are left out.
// Row stride (doubles) of the single-row kernels' power table: slots in
// 16-byte pairs, and an odd number of pairs per row, so that a wave's
// 128-bit reads of different powers fall in different LDS banks.
__host__ __device__ constexpr int pw_stride(int K) { return 2 * (((K + 1) / 2) | 1); }
// Row stride of their radial values f[k][mu]: R rounded up to even (mu pairs in 16 bytes).
__host__ __device__ constexpr int f_stride(int R) { return R + (R & 1); }

__host__ __device__ inline AtomLds2 atom_lds2_layout(int amc, int n_live, int B, int K, int R, int nu_max,
                                                    int n_parts, int nt, int max_partials, int xi_len,
                                                    bool tables = true, int coef_len = 0)
{
    AtomLds2 L;
    const int pw_len = tables ? pw_stride(K) * 3 * (nu_max + 1) : 0;
    const int f_len = f_stride(R) * (K + 1);  // slots padded to an even count
    // even, so that the adjoint region (power table, f) starts 16-byte aligned
    const int mlen = n_live > 0 ? (n_live + 1) & ~1 : 2;
    const int g0 = B > 3 * K ? B : 3 * K;
    int dlen = amc;
    if (tables && dlen < pw_len + f_len) dlen = pw_len + f_len;
    if (dlen < (amc - n_live) + nt) dlen = (amc - n_live) + nt;
    if (dlen < g0 + n_parts * K * 3 + 2 * R * K) dlen = g0 + n_parts * K * 3 + 2 * R * K;
    L.m = kLdsM;
    L.dm = L.m + mlen;
    L.f = L.dm + pw_len;
    L.gp = L.dm + g0;
    L.gf = L.gp + n_parts * K * 3;
    L.gfp = L.gf + R * K;
    L.red = L.dm + dlen - nt;
    L.part = L.dm + dlen;
    L.vx = L.part + (max_partials > 0 ? max_partials : 1);
    L.vy = L.vx + K;
    L.vz = L.vy + K;
    L.r = L.vz + K;
    L.xl = L.r + K;
    L.ints = L.xl + xi_len;
    L.coef = L.ints + (2 * K + nt / kWave + 2) / 2 + 1;
    L.epart = L.coef + coef_len;
    L.total_doubles = L.epart + nt / kWave;
    return L;
}

// LDS of the paired-row kernels: two rows per workgroup whose moments,
// adjoints and partial sums are interleaved in 16-byte slots (row r of slot
// x at double 2x + r), so one contraction entry serves both rows.  Slot 0
// is a zero pair.  Per-row scratch: in phases A/B and after F, row r uses
// the doubles [2 dm + r dlen, + dlen) of the adjoint region; in F, the
// doubles [2 m + r mlen, + mlen) of the dead moment region.
constexpr int kPairM = 1;  // first slot of m

struct AtomLdsPair {
    int m, dm, part;       // first slots (16 bytes each)
    int mlen, dlen;        // slots of the m and dm regions
    int vec;               // double offset of the per-row vectors (vx, vy, vz, r: 4 K per row)
    int ints;              // double offset of the per-row ints (ints_per_row each)
    int ints_per_row;
    int total_doubles;
};

__host__ __device__ inline AtomLdsPair atom_lds_pair_layout(int amc, int n_live, int K, int R, int nu_max, int n_parts,
                                                         int nt, int max_partials)
{
    AtomLdsPair L;
    L.mlen = n_live;
    if (L.mlen < n_parts * K * 3 + 2 * R * K) L.mlen = n_parts * K * 3 + 2 * R * K;
    const int pw_len = (K | 1) * 3 * (nu_max + 1);
    L.dlen = amc;
    if (L.dlen < pw_len + R * K) L.dlen = pw_len + R * K;
    if (L.dlen < (amc - n_live) + nt / 2) L.dlen = (amc - n_live) + nt / 2;
    L.m = kPairM;
    L.dm = L.m + L.mlen;
    L.part = L.dm + L.dlen;
    L.vec = 2 * (L.part + (max_partials > 0 ? max_partials : 1));
    L.ints = L.vec + 8 * K;
    L.ints_per_row = 2 * K + nt / 2 / kWave + 2;
    L.total_doubles = L.ints + (2 * L.ints_per_row + 1) / 2 + 1;
    return L;
}

// Atom-kernel variants (chosen per potential by the launch autotuner):
// v1 = 256 threads, one thread per target moment; v2 = contraction streams
// with 256, 512, 64 or 128 threads per atom (one or two wavefronts: for low
// and mid levels, where a wide workgroup per atom leaves the CU waiting on
// one atom's barriers);
// paired = two atoms per workgroup sharing the contraction streams (plain
// force evaluation only: no per-row coefficients, basis output or row list).
constexpr int kVariantV1 = 0;
constexpr int kVariantV2x256 = 1;
constexpr int kVariantV2x512 = 2;
constexpr int kVariantV2x64 = 3;
constexpr int kVariantPair1024 = 4;  // two rows per 1024-thread workgroup (512 threads per row)
constexpr int kVariantPair512 = 5;   // two rows per 512-thread workgroup
constexpr int kVariantV2x128 = 6;    // streams, 128 threads (two wavefronts per atom, mid levels)
constexpr int kNumVariants = 7;
constexpr bool variant_paired(int v) { return v == kVariantPair1024 || v == kVariantPair512; }
constexpr int variant_threads(int v)
{
    return v == kVariantPair1024 ? 1024
           : ((v == kVariantV2x512 || v == kVariantPair512) ? 512
                                                             : (v == kVariantV2x64 ? 64 : (v == kVariantV2x128 ? 128 : 256)));
}
// gradient parts (thread groups over the basic-moment blocks) per variant
constexpr int variant_parts(int v)
{
    return (v == kVariantV2x512 || v == kVariantPair1024) ? 8 : (v == kVariantV2x64 ? 1 : (v == kVariantV2x128 ? 2 : 4));
}

// global_moments: the moment and adjoint regions in a global slice of
// atom_global_moment_doubles per workgroup instead of LDS (DevPlan::mscratch).
size_t atom_lds_bytes(int amc, int nslots, int R, int nu_max, int n_parts, bool global_moments = false);
size_t atom_global_moment_doubles(int amc, int nslots, int nu_max, int n_parts);
size_t atom_lds2_bytes(int amc, int n_live, int B, int nslots, int R, int nu_max, int n_parts, int nt,
                       int max_partials, int xi_len = 0, bool tables = true, int coef_len = 0);
hipError_t set_atom_kernel_lds(int variant, size_t bytes);
// Resident workgroups per CU of a balanced-item variant's FP64 kernel at
// `bytes` of dynamic LDS (-1 for other variants or on error).
int atom_kernel_blocks_per_cu(int variant, size_t bytes);
// ev_begin / ev_end (may be null) are recorded on `stream` around the atom
// kernel launch, for live per-kernel timing.
// The atom kernel alone over `grid` workgroups (rows, or the persistent
// grid that walks A.rows); the Newton-3 gather of A.g into F.F; the totals.
hipError_t launch_atoms(const DevPlan& P, const AtomArgs& A, size_t lds_bytes, bool radial_fp32, hipStream_t stream,
                        int variant, int grid);
hipError_t launch_gather(const AtomArgs& A, const ForceArgs& F, hipStream_t stream);
hipError_t launch_totals(const AtomArgs& A, const ForceArgs& F, hipStream_t stream);
hipError_t launch_forces(const DevPlan& P, const AtomArgs& A, const ForceArgs& F, size_t lds_bytes,
                         bool radial_fp32, hipStream_t stream, hipEvent_t ev_begin = nullptr,
                         hipEvent_t ev_end = nullptr, int variant = kVariantV1);

constexpr int kAtomBlock = 256;  // threads per atom workgroup; max slots per row

}  // namespace mtp
