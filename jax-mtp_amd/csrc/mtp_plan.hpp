// Evaluation plan: the .mtp alpha tables re-laid for the HIP kernels.
//
// The reference evaluates the level-specific basis either as a full-tensor
// contraction tree (jax_backend.py:162-247) or, in MLIP form, as a
// sequential walk over alpha_index_times (conversion.py:95-102).  The plan
// keeps the MLIP form (same values, ~50x fewer FLOPs at level 24) and
// reorganises it for a GPU:
//
//  * basic moments are grouped into (mu, nu) blocks of complete symmetric
//    tensors in canonical monomial order (verified; every reference level
//    2..28 has this layout);
//  * alpha_index_times is split into dependency layers ("depths"): a target
//    moment's depth is 1 + the max depth of its operands.  Within a layer
//    every target is independent, so one thread owns one target and sums
//    its entries in file order (no atomics, deterministic);
//  * the reverse (adjoint) pass is a gather: for every moment x, the list of
//    (other operand, target, multiplier) of the entries that read x, so that
//    dE/dm[x] = xi(x) + sum mult * m[other] * dE/dm[target], processed from
//    the deepest layer down.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "mtp_format.hpp"

namespace mtp {

struct I4 { int x, y, z, w; };
struct I2 { int x, y; };
struct U2 { uint32_t x, y; };

constexpr int kMaxR = 8;      // radial functions (mu) supported
constexpr int kMaxNu = 12;    // tensor order supported
constexpr int kMaxN = 32;     // Chebyshev terms supported
constexpr int kMaxDepth = 8;  // contraction layers supported

// Device-side view of the plan (all pointers are device pointers).
struct DevPlan {
    int Sp, R, N, B, T, amc, S, nu_max;
    int n_live;                      // moments [0, n_live) are basic or read as operands
    double scaling, min_dist, max_dist;
    const double* radial_coeffs;   // [Sp][Sp][R][N]
    const double* species_coeffs;  // [Sp]
    const double* moment_coeffs;   // [S]
    const int* mapping;            // [S]
    const double* xi_of;           // [amc]  sum of moment_coeffs mapped to each moment
    // basic (mu, nu) blocks
    int blk_off[kMaxR][kMaxNu + 1];  // offset of block (mu, nu) in m[], -1 if absent
    int mu_max_of_nu[kMaxNu + 1];    // largest mu with a (mu, nu) block, -1 if none
    int n_mono;                      // monomials of degree 0..nu_max
    const I4* mono;                // [n_mono] (a, b, c, nu | alpha_in_block << 8)
    // forward layers
    int n_depth;
    int fwd_layer[kMaxDepth + 2];    // target index ranges per depth 1..n_depth (prefix, into fwd_tgt)
    const int* fwd_tgt;              // [n_targets] target moment
    const int* fwd_beg;              // [n_targets + 1] entry ranges
    const uint32_t* fwd_pair;        // [T] i1 | i2 << 16
    const double* fwd_mult;          // [T]
    // reverse layers, deepest first
    int rev_layer[kMaxDepth + 2];    // moment index ranges (prefix, into rev_x)
    const int* rev_x;                // [n_rev] moment
    const int* rev_beg;              // [n_rev + 1]
    const uint32_t* rev_pair;        // [2T] other | target << 16
    const double* rev_mult;          // [2T]
    // per-neighbor gradient: (mu, nu) blocks split into parts
    int n_parts;
    int part_beg[9];                 // prefix into part_blk
    const I2* part_blk;            // (mu, nu)
    // balanced work items for the contraction layers (see build_items).
    // Layer L (forward: depth L+1; reverse: rev_layer index L) owns items
    // [it_beg[L], it_beg[L+1]), combines [cb_beg[L], cb_beg[L+1]), base-only
    // moments [bo_beg[L], bo_beg[L+1]) and a transposed table at tbl_base[L].
    int fwd_it_beg[kMaxDepth + 2], fwd_cb_beg[kMaxDepth + 2], fwd_tbl_base[kMaxDepth + 1];
    int rev_it_beg[kMaxDepth + 2], rev_cb_beg[kMaxDepth + 2], rev_tbl_base[kMaxDepth + 1];
    int rev_bo_beg[kMaxDepth + 2], rev_xa_beg[kMaxDepth + 2];
    const I4* fwd_items;             // (n entries, first moment position, partial slot or -1, -)
    const I4* fwd_comb;              // (moment position, p0, p1, -)
    const U2* fwd_tbl;               // (i1 | i2 << 16, mult | (dest + 1) << 16 on the last entry of a moment)
    const I4* rev_items;
    const I4* rev_comb;
    const U2* rev_tbl;               // (other | target << 16, mult | (dest + 1) << 16)
    const int* rev_bo;               // moment positions without entries (dm = xi)
    const int* rev_xa;               // moments of run items whose dm gets + xi after the sums
    int max_partials;
    // inverse of alpha_moment_mapping: distinct mapped moments (increasing)
    // and, per moment, the scalar indices s mapped to it (increasing)
    int n_mapped;
    const int* map_x;                // [n_mapped]
    const int* map_beg;              // [n_mapped + 1]
    const int* map_s;                // [S]
};

// Work items of one contraction layer for a workgroup of nt threads (see
// split_layer in mtp_plan.cpp).
struct LayerItems {
    std::vector<I4> items, comb;
    std::vector<int> base_only;   // positions
    std::vector<int> xi_add;      // moment ids
    int table_base = 0;
};

struct ItemPlan {
    int nt = 0, max_partials = 0;
    std::vector<LayerItems> fwd, rev;   // fwd[d-1] = depth d; rev[li] = rev_layer li
    std::vector<U2> fwd_table, rev_table;
};

struct HostPlan {
    MtpFile file;           // moments renumbered: never-operand moments last (see build_plan)
    int level = -1;
    int n_live = 0;         // basic moments + moments read as an operand
    int nu_max = 0, n_depth = 0, n_parts = 4;
    std::vector<int> depth;  // [amc]
    std::vector<I4> mono;
    std::vector<int> fwd_tgt, fwd_beg;
    std::vector<uint32_t> fwd_pair;
    std::vector<double> fwd_mult;
    std::vector<int> fwd_layer;
    std::vector<int> rev_x, rev_beg;
    std::vector<uint32_t> rev_pair;
    std::vector<double> rev_mult;
    std::vector<int> rev_layer;
    std::vector<double> xi_of;
    std::vector<I2> part_blk;
    std::vector<int> part_beg;
    std::vector<U2> fwd_pk, rev_pk;
    std::vector<int> map_x, map_beg, map_s;
    int blk_off[kMaxR][kMaxNu + 1];
    int mu_max_of_nu[kMaxNu + 1];
};

void build_items(const HostPlan& p, int nt, ItemPlan& ip);

// Build the plan from a validated file.  n_parts splits the per-neighbor
// gradient work.  Returns false with a message for unsupported layouts.
bool build_plan(const MtpFile& f, int n_parts, HostPlan& p, std::string& err);

// Recompute the parameter-dependent part (xi_of) after a coefficient update.
void refresh_plan_params(HostPlan& p);

}  // namespace mtp
