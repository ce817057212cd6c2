// Launch interface of the training kernels (mtp_train.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mtp_plan.hpp"

namespace mtp {

struct TrainArgs {
    int nrows, nslots, nconf;
    const double* rij;      // [N][K][3]
    const int* itypes;      // [N]
    const int* jtypes;      // [N][K]
    const int* js;          // [N][K] global rows; >= N marks an empty slot
    const int* conf_of;     // [N]
    const double* volumes;  // [C]
    const double* gE;       // [C]
    const double* gF;       // [N][3]
    const double* gS;       // [C][6] reference Voigt order
    double* gxi_part;       // [N][S]
    double* gc_part;        // [N][1 + Sp*R*N]  (alpha, d/dradial for the row's species)
};

constexpr int kTrainThreads = 256;  // threads per row of the training kernel
constexpr int kColsumChunk = 128;   // rows per partial sum of the per-atom gradient columns

// LDS of the training kernel.  A moment travels with its tangent and an
// adjoint with the adjoint of the tangent, as 16-byte pairs, so one entry of
// a contraction stream (the atom kernels' format with slot indices,
// StreamLayout absolute) updates both.  Slots (16 bytes): 0 is a zero pair
// (operand of padding entries, sink of idle lanes); m [n_live] holds
// (m, mt) (never-operand moments alias the adjoint region, dead until the
// reverse pass); dm [dlen] holds (dm, dmt); then the partial sums.  Doubles:
// during the radial/basic-moment phases the adjoint region holds the power
// tables pw, their derivative tables pe and the radial values f, fu; in the
// slot-gradient phase it holds, past the basic moments' adjoints, the
// per-part slot accumulators acc.  Then the per-slot vectors, the
// radial-coefficient factors ca, cb [N][K], a reduction buffer and ints.
struct TrainLds {
    int m, dm, part, dlen;                 // slots
    int pw, pe, f, fu, acc;                // doubles, aliasing the adjoint region
    int vx, vy, vz, r, ex, ey, ez, ud, ca, cb, red, ints, total;  // doubles
};

__host__ __device__ inline TrainLds train_lds_layout(int amc, int n_live, int B, int K, int R, int N, int nu_max,
                                                    int n_parts, int nt, int max_partials)
{
    TrainLds L;
    const int pw_len = (K | 1) * 3 * (nu_max + 1);
    int dlen = amc;
    const int tables = (2 * pw_len + 2 * R * K + 1) / 2;
    if (dlen < tables) dlen = tables;
    if (dlen < B + n_parts * K * R) dlen = B + n_parts * K * R;
    L.m = 1;
    L.dm = L.m + (n_live > 0 ? n_live : 1);
    L.dlen = dlen;
    L.part = L.dm + dlen;
    L.pw = 2 * L.dm;
    L.pe = L.pw + pw_len;
    L.f = L.pe + pw_len;
    L.fu = L.f + R * K;
    L.acc = 2 * (L.dm + B);
    L.vx = 2 * (L.part + (max_partials > 0 ? max_partials : 1));
    L.vy = L.vx + K;
    L.vz = L.vy + K;
    L.r = L.vz + K;
    L.ex = L.r + K;
    L.ey = L.ex + K;
    L.ez = L.ey + K;
    L.ud = L.ez + K;
    L.ca = L.ud + K;
    L.cb = L.ca + N * K;
    L.red = L.cb + N * K;
    L.ints = L.red + nt;
    L.total = L.ints + (K + nt / 64 + 2) / 2 + 1;
    return L;
}

hipError_t set_train_kernel_lds(size_t bytes);
hipError_t launch_train_grad(const DevPlan& P, const TrainArgs& A, size_t lds_bytes, double* scratch_partial,
                             double* xi_sum, double* typed_sum, double* grad, hipStream_t stream);
// Recompute P's xi_of (device) from its moment coefficients (device).
hipError_t launch_xi_of(const DevPlan& P, double* xi, hipStream_t stream);
hipError_t launch_seg_totals(const double* Ei, const double* Wi, const int* conf_off, const double* volumes, int nconf,
                             double* E, double* S6, hipStream_t stream);

}  // namespace mtp
