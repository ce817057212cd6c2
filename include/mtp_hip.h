/*
 * mtp_hip.h — C ABI of the MI355X (gfx950) Moment Tensor Potential evaluator.
 *
 * This library replaces the reference's runtime bridge, which called an
 * exported JAX function through embedded CPython:
 *   ZeroOverhead::ZeroOverheadManager / ZeroOverheadContext
 *     (pair_style_files/zero_overhead.hpp:53-88, zero_overhead.cpp:149-411)
 * and the XLA program it ran:
 *   calc_energy_forces_stress_padded (jax_backend.py:374-460).
 *
 * Plain pointers and sizes only.  All floating point is FP64.  Every entry
 * point returns 0 on success or a negative MTP_ERR_* code; the message is
 * available from mtp_last_error(ctx) (or mtp_global_error() when no context
 * exists yet).  A context is not re-entrant: use one per host thread.
 */
#ifndef MTP_HIP_H
#define MTP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MTP_OK 0
#define MTP_ERR_ARG -1      /* invalid argument or null pointer          */
#define MTP_ERR_FILE -2     /* potential file missing or malformed       */
#define MTP_ERR_CAPACITY -3 /* natoms > max_atoms or count > max_nbr     */
#define MTP_ERR_HIP -4      /* HIP runtime error                         */
#define MTP_ERR_NONFINITE -5 /* NaN/Inf in energy or virial, or in r_ij */
#define MTP_ERR_UNSUPPORTED -6 /* basis layout or size not supported     */

typedef struct mtp_ctx mtp_ctx;

/* Open a trained .mtp potential and size device buffers for at most
 * max_atoms atoms of max_nbr neighbor slots each.
 * Replaces PairJaxMTP::settings/init_style + ZeroOverheadContext
 * (pair_jax_mtp.cpp:228-245, 264-282; zero_overhead.cpp:149-169, 408-411).  The first pair_style argument is
 * the .mtp file (the reference took a jax.export .bin).  cutoff is the
 * MTP_CUTOFF argument (used by the packing helper); debug_level is kept. */
int mtp_open(const char *mtp_path, int max_atoms, int max_nbr, double cutoff, int debug_level, mtp_ctx **out);

/* Same, from the text of a .mtp file held in memory. */
int mtp_open_text(const char *mtp_text, int max_atoms, int max_nbr, double cutoff, int debug_level, mtp_ctx **out);

/* Release the context and its device memory. */
void mtp_close(mtp_ctx *ctx);

/* Last error message of ctx ("" if none); mtp_global_error() for errors
 * raised before a context existed (e.g. by mtp_open). */
const char *mtp_last_error(const mtp_ctx *ctx);
const char *mtp_global_error(void);

/* Radial-basis precision: 0 = FP64 (default), 1 = FP32 radial basis with
 * FP64 moments, contraction, energy and virial. */
int mtp_set_precision(mtp_ctx *ctx, int radial_fp32);

/* Basis description: level (conversion.py:14-29), species count, R, N,
 * alpha_moments_count, B, T, S.  Any pointer may be NULL. */
int mtp_info(const mtp_ctx *ctx, int *level, int *species_count, int *radial_funcs_count, int *radial_basis_size,
             int *alpha_moments_count, int *basic_count, int *times_count, int *scalar_count);

/* Replace the trainable parameters in place (moment_coeffs[S] in .mtp
 * order, species_coeffs[Sp], radial_coeffs[Sp][Sp][R][N]); NULL keeps a
 * group.  Mirrors MTPData.parameters' setter (data.py:74-102). */
int mtp_set_params(mtp_ctx *ctx, const double *moment_coeffs, const double *species_coeffs,
                   const double *radial_coeffs, double scaling); /* scaling NaN or <= 0 keeps the current one */

/* Same from device arrays, enqueued on stream (no host synchronisation,
 * so it can be captured in a graph with the evaluation that follows).  The
 * host copy of the parameters (mtp_set_params, mtp_info) is not updated.
 * No reference counterpart: the reference re-traced its jitted loss with new
 * parameters (motep_jax_train_function.py:154-191). */
int mtp_set_params_device(mtp_ctx *ctx, const double *d_moment_coeffs, const double *d_species_coeffs,
                          const double *d_radial_coeffs, void *stream);

/* Host-array evaluation with the data contract of
 * ZeroOverheadManager::execute_potential (zero_overhead.hpp:58-64), as the
 * pair style calls it (pair_jax_mtp.cpp:183-189):
 * rij_rows[i] holds 3*max_nbr doubles (slot k at 3k..3k+2), nbr_rows[i] and
 * nbr_type_rows[i] max_nbr ints, nbr_counts[i] the valid prefix; rows exist
 * for i < natoms.  nneigh_total is accepted and unused (as in the reference).
 * Outputs: energy (sum of E_i, i < natoms), forces[natoms*3], stress6 in the
 * reference convention (jax_backend.py:357-360: (W+W^T)/(2V), order xx yy zz
 * yz xz xy), virial6 in the LAMMPS convention (-sym(W), order xx yy zz xy xz
 * yz, energy units).  stress6/virial6 may be NULL.  Synchronous. */
int mtp_compute(mtp_ctx *ctx, int natoms, int nneigh_total, const double *const *rij_rows, const int *types,
                const int *const *nbr_rows, const int *nbr_counts, const int *const *nbr_type_rows, double volume,
                double *energy, double *forces, double *stress6, double *virial6);

/* The pair style's whole force call with the packing and staging on the
 * GPU: the LAMMPS arrays as they are (x[nall][3] contiguous, 1-based
 * type[nall], tag[nall], the full neighbor list ilist[inum],
 * numneigh[nlocal], firstneigh[nlocal], indices may carry LAMMPS's special
 * bits) and the orthogonal box (prd[3], periodic[3]).  The host flattens
 * the list into pinned memory and uploads it with the atoms; a kernel then
 * does the walk of PairJaxMTP::compute (pair_jax_mtp.cpp:96-173: minimum
 * image, r^2 < MTP_CUTOFF^2 while fewer than MAX_ATOMS_PER_NEIGHBOR are
 * kept, ghosts folded onto owners by tag, unmapped -> nlocal / type -1) and
 * the staging of zero_overhead.cpp:61-117, bit for bit what
 * mtp_pack_full_list + mtp_compute produce, and the evaluation follows.
 * Replaces that pair (pair_jax_mtp.cpp:96-189) for the pair style.
 * list_is_new = 0 promises that the list, ilist, type and tag are those of
 * the previous call on ctx (LAMMPS: neighbor->ago > 0, no rebuild since):
 * only x is uploaded then.  A change of nlocal, nall, inum or the list
 * length uploads everything regardless.
 * Outputs as mtp_compute; *n_truncated (may be NULL) receives the in-cutoff
 * neighbors dropped by the row limit.  mtp_last_atom_values works after it.
 * Synchronous. */
int mtp_compute_full_list(mtp_ctx *ctx, int nlocal, int nall, const double *x, const int *type, const int *tag,
                          int inum, const int *ilist, const int *numneigh, const int *const *firstneigh,
                          int list_is_new, const double *prd, const int *periodic, double volume, double *energy,
                          double *forces, double *stress6, double *virial6, long long *n_truncated);

/* Device-resident evaluation with the 8 dynamic arguments of
 * calc_energy_forces_stress_padded (jax_backend.py:374-392; dtypes as
 * compile_cuda.py:225-232 but FP64): d_itypes[nrows], d_js[nrows*nslots],
 * d_rijs[nrows*nslots*3], d_jtypes[nrows*nslots], cell_rank, volume,
 * natoms_energy, natoms_force.  nrows <= max_atoms, nslots <= max_nbr.
 * Masks as the reference wires them: the energy sums E_i for
 * i < natoms_energy (jax_backend.py:447-452), and the forces are masked by
 * natoms_energy too (rows i and Newton-3 targets j below it), because
 * jax_backend.py:414-418 passes natoms_energy into the natoms_force slot of
 * calc_energy_forces_stress; natoms_force is accepted and unused.
 * Device outputs: d_out[14] = {E, stress6 (reference), virial6 (LAMMPS),
 * status}, d_forces[nrows*3]; d_eatom[nrows] (per-atom energies) may be
 * NULL.  status bit 1: non-finite energy or virial; bit 2: a slot vector
 * r_ij with a NaN/Inf component (the slot is dropped and E set to NaN).
 * stream is a hipStream_t (NULL = the context's stream).  Returns after
 * enqueueing; call mtp_sync(ctx, stream) before reading results. */
int mtp_compute_padded_device(mtp_ctx *ctx, int nrows, int nslots, const int *d_itypes, const int *d_js,
                              const double *d_rijs, const int *d_jtypes, int cell_rank, double volume,
                              int natoms_energy, int natoms_force, double *d_out, double *d_forces,
                              double *d_eatom, void *stream);

/* Per-atom energies and virials of the last mtp_compute on ctx, for the pair
 * style's eflag_atom / vflag_atom (no reference counterpart: the reference
 * returns totals only, pair_jax_mtp.cpp:195-205).  eatom[natoms] = E_i;
 * vatom6[natoms][6] = -(W_i + W_i^T) / 2 with W_i = sum_k r_ik (x) dE_i/dr_ik
 * (atom i's row), in LAMMPS order (xx, yy, zz, xy, xz, yz, energy units), so
 * that the vatom6 rows sum to mtp_compute's virial6.  Either output may be
 * NULL.  natoms must not exceed that call's natoms (MTP_ERR_ARG otherwise,
 * also after a device entry point has run on ctx since).  Synchronous. */
int mtp_last_atom_values(mtp_ctx *ctx, int natoms, double *eatom, double *vatom6);

/* Device pointer of the per-slot gradients dE_i/dr_ik [nrows][nslots][3] of
 * the last device evaluation (valid until the next call on ctx). */
const double *mtp_last_slot_gradients(const mtp_ctx *ctx);

/* Wait for work enqueued on stream (NULL = the context's stream). */
int mtp_sync(mtp_ctx *ctx, void *stream);

/* Atom-kernel variants (no reference counterpart; XLA chose its own
 * fusion): 0 = one thread per target moment, 256 threads per atom;
 * 1 / 2 / 3 = contraction streams with 256 / 512 / 64 threads per atom;
 * 4 / 5 = two atoms per workgroup sharing the contraction streams, 1024 /
 * 512 threads (plain force evaluation only); 6 = contraction streams with
 * 128 threads per atom.  mtp_open picks variant 3 for small contraction
 * tables (<= 400 entries, levels up to 14), variant 6 up to 8,000 entries
 * (levels up to 20) and otherwise the widest single-row variant whose LDS
 * fits (env MTP_KERNEL_VARIANT overrides).  mtp_autotune times every available variant on the given
 * device arrays (reps launches each, after one warm-up), selects the
 * fastest for ctx and reports the per-variant ms in
 * ms_per_variant[MTP_NUM_VARIANTS] (-1 = unavailable). */
#define MTP_NUM_VARIANTS 7
int mtp_set_variant(mtp_ctx *ctx, int variant);
int mtp_get_variant(const mtp_ctx *ctx, int *variant, long long *lds_bytes);
int mtp_autotune(mtp_ctx *ctx, int nrows, int nslots, const int *d_itypes, const int *d_js, const double *d_rijs,
                 const int *d_jtypes, int reps, int *best, double *ms_per_variant);

/* Live per-kernel timing (no reference counterpart; the reference only
 * timed whole calls, compile_cuda.py:252-275).  While enabled, HIP events
 * are recorded on the launch stream around every atom-kernel launch;
 * mtp_profile_read returns the summed atom-kernel time (ms) and the launch
 * count since the last read, then resets. */
int mtp_profile(mtp_ctx *ctx, int enable);
int mtp_profile_read(mtp_ctx *ctx, double *atom_kernel_ms, int *launches);

/* ---- training (packed batches; device arrays) ------------------------
 * A batch of nconf configurations is nrows concatenated atom rows;
 * d_conf_off[nconf+1] gives each configuration's rows, d_conf_of[nrows] the
 * configuration of each row.  d_js holds global row indices (>= nrows for an
 * empty slot).  Replaces the reference's global padding of every
 * configuration to (A_max, K_max) (motep_jax_train_import.py:219-295). */

/* Per-configuration energies d_E[nconf], forces d_forces[nrows*3] and
 * stress d_stress[nconf*6] (reference Voigt order, / V_c)
 * (jax.py:29-98 semantics with the Newton-3 scatter of jax_backend.py:263-303). */
int mtp_batch_forward(mtp_ctx *ctx, int nrows, int nslots, int nconf, const int *d_conf_off, const int *d_conf_of,
                      const int *d_itypes, const int *d_js, const double *d_rijs, const int *d_jtypes,
                      const double *d_volumes, double *d_E, double *d_forces, double *d_stress, void *stream);

/* Hand-written parameter gradient of sum gE.E + sum gF.F + sum gS.S for
 * upstream gradients d_gE[nconf], d_gF[nrows*3], d_gS[nconf*6], written to
 * d_grad[S + Sp + Sp*Sp*R*N] in the order (moment_coeffs, species_coeffs,
 * radial_coeffs) of MTPData.parameters (data.py:60-72).  Replaces JAX's
 * value_and_grad of the training loss (motep_jax_train_function.py:191). */
int mtp_batch_param_grad(mtp_ctx *ctx, int nrows, int nslots, int nconf, const int *d_conf_off, const int *d_conf_of,
                         const int *d_itypes, const int *d_js, const double *d_rijs, const int *d_jtypes,
                         const double *d_volumes, const double *d_gE, const double *d_gF, const double *d_gS,
                         double *d_grad, void *stream);

/* ---- basis values, per-row coefficients, basis Jacobians -------------
 * No counterpart in the reference (it has no learned head).  They expose the
 * pieces of the atom kernel the distilled head is built from.  All arrays
 * are device arrays; rows and slots as in mtp_compute_padded_device. */

/* Basis values b[row][s] = m[alpha_moment_mapping[s]] (the scalar moments
 * that E_i = species_coeff + moment_coeffs . b_i contracts; forward pass
 * only). */
int mtp_descriptor_device(mtp_ctx *ctx, int nrows, int nslots, const int *d_itypes, const double *d_rijs,
                          const int *d_jtypes, double *d_basis, void *stream);

/* Slot gradients d_g[row][slot][3] = d(xi_row . b_row)/d r_row,slot and
 * per-row virials d_Wi[row][9] (may be NULL) for per-row coefficient vectors
 * xi_row = d_xi + row * xi_stride (xi_stride 0: one vector for all rows). */
int mtp_slot_grad_xi_device(mtp_ctx *ctx, int nrows, int nslots, const int *d_itypes, const double *d_rijs,
                            const int *d_jtypes, const double *d_xi, int xi_stride, double *d_g, double *d_Wi,
                            void *stream);

/* Dense basis Jacobian d_J[row][s][slot][3] = d b_row,s / d r_row,slot. */
int mtp_basis_jacobian_device(mtp_ctx *ctx, int nrows, int nslots, const int *d_itypes, const double *d_rijs,
                              const int *d_jtypes, double *d_J, void *stream);

/* d_g[row][slot] = sum_s d_xi[row][s] d_J[row][s][slot] */
int mtp_jacobian_apply_device(int nrows, int nslots, int S, const double *d_J, const double *d_xi, double *d_g,
                              void *stream);

/* d_v[row][s] = [row < n_f] sum_slot d_J[row][s][slot] . (gF[row] - [j < n_f] gF[j]), j = d_js[row][slot]:
 * the adjoint of the Newton-3 force gather applied to force gradients gF. */
int mtp_jacobian_tapply_device(int nrows, int nslots, int S, int n_f, const double *d_J, const int *d_js,
                               const double *d_gF, double *d_v, void *stream);

/* Newton-3 force gather of slot gradients d_g (jax_backend.py:263-303) with
 * the context's scratch. */
int mtp_gather_forces_device(mtp_ctx *ctx, int nrows, int nslots, const int *d_js, const double *d_g, int n_f,
                             double *d_forces, void *stream);

/* The same gather without a context and without allocating: for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).  No counterpart in the
 * reference.  The reverse-neighbor CSR lives in a static per-device scratch
 * array of the library (18.4 MB per GPU, held until the process ends), so
 * nrows <= 131072 and nrows * nslots <= 4194304 (MTP_ERR_CAPACITY beyond).
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one
 * GPU (other streams, other threads) race.  The CSR is rebuilt on every
 * call (no reuse across calls).  Errors: mtp_global_error().  Enqueued on
 * stream (NULL = the null stream); results are ready after it completes. */
int mtp_gather_forces_noalloc(int nrows, int nslots, const int *d_js, const double *d_g, int n_f, double *d_forces,
                              void *stream);

/* ---- distilled head ---------------------------------------------------
 * No counterpart in the reference.  An MLP over the basis values of a
 * low-level potential `desc` (level 12: 29 values) stands in for a
 * high-level potential `exact`, which re-evaluates every atom whose
 * normalised basis values leave the training box [lo, hi] (per species).
 * Network: z = [(b - mean) * inv_std, onehot(type)] (S + Sp <= 32),
 * 64 SiLU, 64 SiLU, E_i = e_shift[type] + e_scale * out; with residual != 0
 * the descriptor potential's own energy species_coeff + moment_coeffs . b
 * is added (the network corrects the low-level potential; its current
 * coefficients are used).  Parameters theta
 * (mtp_head_theta_size() doubles): W1[32][64] b1[64] W2[64][64] b2[64]
 * w3[64] b3, weights stored (in, out). */
typedef struct mtp_head mtp_head;

int mtp_head_theta_size(void);
/* exact may be NULL (no fallback); lo/hi [Sp][S] may both be NULL (no box).
 * Both potentials must stay open while the head is. */
int mtp_head_open(mtp_ctx *desc, mtp_ctx *exact, const double *theta, const double *mean, const double *inv_std,
                  const double *lo, const double *hi, const double *e_shift, double e_scale, int residual,
                  mtp_head **out);
void mtp_head_close(mtp_head *h);
const char *mtp_head_last_error(const mtp_head *h);
/* theta from host (on_device 0, synchronous) or device memory (stream-ordered). */
int mtp_head_set_theta(mtp_head *h, const double *theta, int on_device, void *stream);

/* Same outputs and masks as mtp_compute_padded_device (d_out[14]: E,
 * stress6, virial6, status; n_f accepted and unused), plus d_flags[nrows] (1 = out of box; may be NULL) and
 * d_n_fallback (device int: rows re-evaluated exactly; may be NULL). */
int mtp_head_compute_padded_device(mtp_head *h, int nrows, int nslots, const int *d_itypes, const int *d_js,
                                   const double *d_rijs, const int *d_jtypes, int cell_rank, double volume, int n_e,
                                   int n_f, double *d_out, double *d_forces, double *d_eatom, int *d_flags,
                                   int *d_n_fallback, void *stream);

/* Network alone on given basis values: d_E[row], d_xi[row][s] = dE_row/db_row,s. */
int mtp_head_eval_device(mtp_head *h, int nrows, const int *d_itypes, const double *d_basis, double *d_E,
                         double *d_xi, int *d_flags, void *stream);
/* d_grad[theta] = d/dtheta sum_row [alpha_row E_row + v_row . dE_row/db_row]
 * (the head's training gradient: alpha = dL/dE, v = J^T dL/dg). */
int mtp_head_param_grad_device(mtp_head *h, int nrows, const int *d_itypes, const double *d_basis,
                               const double *d_alpha, const double *d_v, double *d_grad, void *stream);

/* ---- host-only helpers (no GPU work; usable without a device) -------- */

/* Parse, validate and plan a .mtp file without touching the GPU.  Returns
 * the same fields as mtp_info plus the LDS bytes the atom kernel needs for
 * max_nbr slots. */
int mtp_check_file(const char *mtp_path, int max_nbr, int *level, int *species_count, int *alpha_moments_count,
                   int *basic_count, int *times_count, int *scalar_count, long long *lds_bytes);

/* Neighbor packing of PairJaxMTP::compute (pair_jax_mtp.cpp:96-173) for a
 * full neighbor list and an orthogonal box:
 * pos[max_atoms][3*max_nbr], nbr/ntype[max_atoms][max_nbr], counts[max_atoms],
 * atom_types[max_atoms].  Returns nneigh_total (>= 0) or MTP_ERR_CAPACITY
 * when nlocal > max_atoms.  Neighbor overflow truncates like the reference;
 * *n_truncated (may be NULL) receives the number of dropped in-cutoff
 * neighbors. */
long long mtp_pack_full_list(int nlocal, const double *x, const int *type, const int *tag, int inum,
                             const int *ilist, const int *numneigh, const int *const *firstneigh, double cutoff,
                             const double *prd, const int *periodic, int max_atoms, int max_nbr, double *pos,
                             int *nbr, int *ntype, int *counts, int *atom_types, long long *n_truncated);

/* Staging of zero_overhead.cpp:61-117 (update_atom_data_zero_copy) in FP64:
 * padded [rows][max_nbr] arrays with sentinels j = natoms, type = -1,
 * r = 2000, itype = -1, then the valid prefix of each row i < natoms. */
int mtp_stage_host(int rows, int max_nbr, int natoms, const double *const *rij_rows, const int *types,
                   const int *const *nbr_rows, const int *nbr_counts, const int *const *nbr_type_rows,
                   int *itypes_out, int *js_out, double *rij_out, int *jtypes_out);

#ifdef __cplusplus
}
#endif
#endif /* MTP_HIP_H */
