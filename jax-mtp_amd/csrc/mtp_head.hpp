// Distilled head: an MLP over a low-level MTP's basis values (the
// "invariants" b_i of atom i) that replaces the level-24 contraction tree,
// with the exact potential as fallback for atoms outside the training box.
//
// Network (one per potential, shared by all species; the species enters as
// a one-hot input):
//   z_i  = [(b_i - mean) * inv_std, onehot(t_i)]             (S + Sp <= 64)
//   a1   = W1^T z + b1,  h1 = silu(a1)                        (64)
//   a2   = W2^T h1 + b2, h2 = silu(a2)                        (64)
//   E_i  = e_shift[t_i] + e_scale * (w3 . h2 + b3)
//          [+ lin_sc[t_i] + lin . b_i   in residual form: the network then
//           corrects the low-level potential's own linear energy]
// Flat parameter vector theta (weights stored (in, out), row major):
//   W1[64][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
// Rows of W1 beyond S + Sp multiply zero inputs.
#pragma once
#include <hip/hip_runtime.h>

namespace mtp {

constexpr int kHeadHidden = 64;  // = wavefront width: lane j owns hidden unit j
constexpr int kHeadIn = 64;      // padded input width (= wavefront width: lane i owns input i)
constexpr int kHeadTheta = kHeadIn * kHeadHidden + kHeadHidden + kHeadHidden * kHeadHidden + 2 * kHeadHidden + 1;

struct HeadNet {
    int S, Sp;
    const double* theta;    // [kHeadTheta] canonical
    const double* wt;       // [kHeadHidden][kHeadIn] W1^T then [kHeadHidden][kHeadHidden] W2^T (prepared)
    const double* mean;     // [S]
    const double* inv_std;  // [S]
    const double* lo;       // [Sp][S] box of the training z values
    const double* hi;       // [Sp][S]
    const double* e_shift;  // [Sp]
    double e_scale;
    const double* lin;      // [S] linear coefficients (residual form) or null
    const double* lin_sc;   // [Sp] species energies of the linear part
};

// theta -> transposed copies used by the backward passes.
hipError_t head_prepare(const HeadNet& H, double* wt, hipStream_t s);

// E_i, dE_i/db_i and the out-of-box flag of rows [0, n).  flags/count may be
// null; with count, flagged rows (itype >= 0 only) are appended to `rows`
// (order unspecified; *count must be zero on entry: head_zero_count).
// exact_mask (may be null): per row 1 if the row is appended to rows (its
// inputs leave the training box and its type is valid), else 0.
hipError_t head_eval(const HeadNet& H, int n, const int* itypes, const double* basis, double* E, double* xi,
                     int* flags, int* rows, int* count, hipStream_t s, int* exact_mask = nullptr);
hipError_t head_zero_count(int* count, hipStream_t s);

// d/dtheta of sum_i [alpha_i E_i + v_i . dE_i/db_i]  (v_i in basis units).
// part must hold head_grad_partials(n) doubles.
size_t head_grad_partials(int n);
hipError_t head_param_grad(const HeadNet& H, int n, const int* itypes, const double* basis, const double* alpha,
                           const double* v, double* part, double* grad, hipStream_t s);

// Basis Jacobians J[A][S][K][3] (dense, per row):
//   g[a][k] = sum_s xi[a][s] J[a][s][k]
//   v[a][s] = [a < n_f] sum_k J[a][s][k] . (gF[a] - [js[a][k] < n_f] gF[js[a][k]])
// (the adjoint of the force gather F_a = [a < n_f] sum_k g_ak
//  - sum_{i < n_f, k: js[i][k] = a < n_f} g_ik).
hipError_t jacobian_apply(int n, int K, int S, const double* J, const double* xi, double* g, hipStream_t s);
hipError_t jacobian_tapply(int n, int K, int S, int n_f, const double* J, const int* js, const double* gF, double* v,
                           hipStream_t s);

}  // namespace mtp
