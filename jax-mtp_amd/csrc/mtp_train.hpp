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

struct TrainLds {
    int m, mt, dm, dmt, f, fp, vx, vy, vz, r, dx, dy, dz, ud, a, b, red, ints, total;
};

size_t train_lds_bytes(int amc, int K, int R, int N, int nu_max, int n_parts);
hipError_t set_train_kernel_lds(size_t bytes);
hipError_t launch_train_grad(const DevPlan& P, const TrainArgs& A, size_t lds_bytes, double* scratch_partial,
                             double* xi_sum, double* typed_sum, double* grad, hipStream_t stream);
// Recompute P's xi_of (device) from its moment coefficients (device).
hipError_t launch_xi_of(const DevPlan& P, double* xi, hipStream_t stream);
hipError_t launch_seg_totals(const double* Ei, const double* Wi, const int* conf_off, const double* volumes, int nconf,
                             double* E, double* S6, hipStream_t stream);

}  // namespace mtp
