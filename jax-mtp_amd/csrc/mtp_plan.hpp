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
struct U4 { uint32_t x, y, z, w; };

constexpr int kWave = 64;     // wavefront width of gfx950
constexpr int kMaxR = 8;      // radial functions (mu) supported
constexpr int kMaxNu = 12;    // tensor order supported
// Shapes of the monomial-column tasks of the basic moments: NMU mu per
// monomial, C monomials (NMU * C <= 8 accumulators).
constexpr int kColShapes = 5;
constexpr int kColNmu[kColShapes] = {1, 2, 4, 6, 8};
constexpr int kColLen[kColShapes] = {8, 4, 2, 1, 1};
constexpr int kColTasksPerWave = 16;  // four lanes per task
constexpr int kMaxN = 32;     // Chebyshev terms supported
constexpr int kMaxDepth = 8;  // contraction layers supported
constexpr int kStreamUnroll = 3;  // steps per prefetch batch of a contraction stream
// Multiplier word of a contraction-stream step (see DevPlan): bits 0-17
// the item's destination as an LDS byte offset from the LDS base on the last
// step of the lane's item, else 0 (the zero slot at offset 0 is never a
// destination, so a nonzero field is the item's end); bits 18-31 the
// entries' integer multiplier (0 in padding steps).  The kernel takes the
// store address with one AND and the multiplier with one shift.
constexpr uint32_t kStreamDestMask = 0x3ffffu;
constexpr int kStreamMultShift = 18;
constexpr uint32_t kStreamMultMax = 0x3fffu;
constexpr uint32_t stream_mult(uint32_t w) { return w >> kStreamMultShift; }
constexpr uint32_t stream_dest(uint32_t w) { return w & kStreamDestMask; }

constexpr int kLdsM = 2;      // m[] starts at this double of the atom kernels' LDS (lds[0] is a zero)

// Device-side view of the plan (all pointers are device pointers).
struct DevPlan {
    int Sp, R, N, B, T, amc, S, nu_max;
    int n_live;                      // moments [0, n_live) are basic or read as operands
    int zero_untargeted;             // 1: some non-basic moment is no entry's target (it stays 0)
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
    const I4* mono;                // [n_mono] (a, b, c, nu | alpha_in_block << 8 | wave's max mu << 24)
    const I4* mono_dst;            // [n_mono][2]: basic moment index of (mu, monomial), mu = 0..7, or -1
    int n_cwaves;                    // wavefronts of monomial-column tasks (kColTasksPerWave tasks each)
    const I4* ctask;               // [n_cwaves * kColTasksPerWave][3]: (a, b, c0, shape), 8 destinations
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
    const I2* part_blk;            // (mu | offset of the (mu, nu) block << 8, nu)
    // contraction layers as per-wavefront entry streams (see build_items).
    // Layer L (forward: depth L+1; reverse: rev_layer index L) gives wave w
    // of the workgroup the steps [wave[L * NW + w].x, + .y) of tbl (.y a
    // multiple of kStreamUnroll).  Step s holds up to three entries per lane
    // that share a multiplier: tbl[s * 64 + lane] = (e0, e1, e2, w) with
    // e = a | b << 16, a and b LDS byte offsets of the operands (a from the
    // workgroup's LDS base; b from it in the forward pass, from the adjoint
    // region in the reverse pass; 0 for an unused entry: the zero slot) and
    // w the multiplier word (stream_mult), so that the lane's sum gains
    // mult (a0 b0 + a1 b1 + a2 b2).  Each lane walks its own sequence of
    // items: on the last step of an item (stream_dest(w) nonzero, per lane)
    // the lane stores its sum at that LDS byte offset and restarts from zero.
    // Rows split into chunks leave partial sums, combined per layer by
    // comb[cb_beg[L] .. cb_beg[L+1]] = (moment, first partial, end partial,
    // -) in a fixed order.
    int fwd_cb_beg[kMaxDepth + 2], rev_cb_beg[kMaxDepth + 2];
    int rev_bo_beg[kMaxDepth + 2], rev_xa_beg[kMaxDepth + 2];
    int abs_slots;                   // 1: operand fields are 8-byte slot indices from the LDS base (a
                                     // and b, both directions; StreamLayout absolute), for layouts whose
                                     // byte offsets do not fit 16 bits; 0: byte offsets as above
    const I4* fwd_wave;
    const U4* fwd_tbl;
    const I4* fwd_comb;
    const I4* rev_wave;
    const U4* rev_tbl;
    const I4* rev_comb;
    const int* rev_bo;               // moments without entries (dm = xi)
    const int* rev_xa;               // moments of unsplit rows whose dm gets + xi after the sums
    int k_layout;                    // slots the LDS layout is sized for (>= the slots of any call)
    int lds_tables;                  // 1: the LDS layout holds the power table and radial values; 0 only
                                     // for the plans of a per-row pass that reads a moment cache
    int coef_lds;                    // > 0: the radial coefficients of the row's species (Sp * R * N
                                     // doubles) are staged in LDS; 0: read from global memory
    int max_partials;
    // Per-target kernel with the moments and adjoints in global memory (levels
    // whose moments exceed the LDS): a grid of mscratch_grid workgroups walks
    // the rows, each workgroup in its slice of mscratch_stride doubles.  Null
    // when they live in LDS.
    double* mscratch;
    int mscratch_grid;
    long long mscratch_stride;
    // inverse of alpha_moment_mapping: distinct mapped moments (increasing)
    // and, per moment, the scalar indices s mapped to it (increasing)
    int n_mapped;
    const int* map_x;                // [n_mapped]
    const int* map_beg;              // [n_mapped + 1]
    const int* map_s;                // [S]
};

// Contraction streams of one direction (all layers) for a workgroup of nt
// threads, in the DevPlan layout above.
struct StreamPlan {
    std::vector<I4> wave;          // [n_layers * nt / 64]
    std::vector<U4> tbl;           // [steps * 64]
    std::vector<I4> comb;
    std::vector<int> bo, xa;
    int cb_beg[kMaxDepth + 2], bo_beg[kMaxDepth + 2], xa_beg[kMaxDepth + 2];
    int max_partials = 0;
    double cost = 0;               // modelled steps of the slowest wave, summed over layers
};

struct ItemPlan {
    int nt = 0, max_partials = 0;
    StreamPlan fwd, rev;
};

struct HostPlan {
    MtpFile file;           // moments renumbered: never-operand moments last (see build_plan)
    int level = -1;
    int n_live = 0;         // basic moments + moments read as an operand
    int nu_max = 0, n_depth = 0, n_parts = 4;
    std::vector<int> depth;  // [amc]
    std::vector<I4> mono;
    std::vector<I4> mono_dst;   // [n_mono * 2]
    std::vector<I4> ctask;      // [3 per task] monomial-column tasks of the basic moments
    int n_cwaves = 0;
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
    std::vector<int> map_x, map_beg, map_s;
    int blk_off[kMaxR][kMaxNu + 1];
    int mu_max_of_nu[kMaxNu + 1];
};

// Where the contraction streams find their operands.  Moments, adjoints and
// partial sums occupy slots of `unit` bytes: 8 for one row per workgroup,
// 16 for two rows whose values are interleaved (both rows share every
// entry).  Offsets are in slots from the LDS base.  With `absolute`, the
// operand fields hold slot indices from the LDS base; otherwise they hold
// byte offsets (a and the forward b from the LDS base, the reverse b from
// dm_off).  Item destinations are byte offsets from the LDS base.
struct StreamLayout {
    int unit = 8;
    int m_off = kLdsM;
    int dm_off = 0;
    int part_off = 0;
    bool absolute = false;
};

// Streams for a workgroup of nt threads.  part_budget caps the partial sums
// of a layer when the chunking allows.  Returns false when an operand
// offset does not fit the 16-bit fields, a destination the 18-bit field,
// or a multiplier is not an integer in [1, kStreamMultMax].
bool build_items(const HostPlan& p, int nt, const StreamLayout& lay, int part_budget, ItemPlan& ip,
                 std::string& err);

// Build the plan from a validated file.  n_parts splits the per-neighbor
// gradient work.  Returns false with a message for unsupported layouts.
bool build_plan(const MtpFile& f, int n_parts, HostPlan& p, std::string& err);

// Recompute the parameter-dependent part (xi_of) after a coefficient update.
void refresh_plan_params(HostPlan& p);

}  // namespace mtp
