/*
 * FP64 CPU oracle for the MTP evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The shipped library (libmtp_hip.so) never links or calls it.
 *
 * It restates, in plain C and FP64, the reference algorithm of
 * sebastian57/JAX-MTP for one force evaluation:
 *
 *   radial basis      jax_backend.py:83-95, 108-160
 *   compact moments   conversion.py:39-102 (MLIP alpha tables)
 *   full-tensor tree  jax_backend.py:162-247 (JSON contraction tree,
 *                     tensordot per moment_jax.py:49-74)
 *   per-atom gradient jax_backend.py:20-57 (value_and_grad, here a
 *                     hand-derived reverse pass, independent per form)
 *   force scatter     jax_backend.py:263-303
 *   stress            jax_backend.py:355-370
 *   energy mask       jax_backend.py:431-456
 *   neighbor packing  pair_jax_mtp.cpp:96-173 and zero_overhead.cpp:61-117
 *
 * Two independent evaluators of E_i and g_ik = dE_i/dr_ik are provided:
 * the compact MLIP-table form and the full-tensor tree form.  They share
 * only the radial basis.  Tests require them to agree to ~1e-13.
 *
 * Deviation kept on purpose (documented in the tests): slots that are not
 * "valid" (type < 0 or r >= max_dist) are skipped.  The reference multiplies
 * them by a zero smoothing factor, which gives the same value whenever the
 * slot's r is non-zero and finite, as the reference's own sentinels are
 * (20.0 in pair_jax_mtp.cpp:160-166, 2000.0f in zero_overhead.cpp:78).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int Sp, R, N, B, T, amc, S;
    double scaling, min_dist, max_dist;
    const double *radial_coeffs;  /* [Sp][Sp][R][N] */
    const double *species_coeffs; /* [Sp] */
    const double *moment_coeffs;  /* [S]  .mtp (file) order */
    const int *basic;             /* [B][4]  (mu, a, b, c) */
    const int *times;             /* [T][4]  (i1, i2, mult, i3) */
    const int *mapping;           /* [S] */
} orc_mtp;

/* jnp gather semantics: negative indices count from the end, out-of-range
 * indices clamp (jax_backend.py:94 and :101 index with itype / jtypes). */
static int wrap_species(int t, int Sp)
{
    if (t < 0) t += Sp;
    if (t < 0) t = 0;
    if (t >= Sp) t = Sp - 1;
    return t;
}

/* Chebyshev basis T_0..T_{N-1} and dT/dx (jax_backend.py:108-144). */
static void chebyshev(int N, double x, double *T, double *dT)
{
    if (N <= 0) return;
    T[0] = 1.0; dT[0] = 0.0;
    if (N == 1) return;
    T[1] = x; dT[1] = 1.0;
    for (int n = 2; n < N; ++n) {
        T[n] = 2.0 * x * T[n - 1] - T[n - 2];
        dT[n] = 2.0 * T[n - 1] + 2.0 * x * dT[n - 1] - dT[n - 2];
    }
}

/* Radial functions f_mu(r) and df_mu/dr for one slot.  Returns 1 when the
 * slot is valid (jax_backend.py:86-95, 146-160), else 0 with f = df = 0.
 * If Tn is given it receives scaling*s*T_n (the factor multiplying
 * c[ti][tk][mu][n]). */
static int radial(const orc_mtp *P, double r, int ti, int tk, double *f, double *df, double *Tn)
{
    int R = P->R, N = P->N;
    int valid = (tk >= 0) && (r < P->max_dist);
    if (!valid) {
        for (int m = 0; m < R; ++m) f[m] = df[m] = 0.0;
        if (Tn) for (int n = 0; n < N; ++n) Tn[n] = 0.0;
        return 0;
    }
    double T[64], dT[64];
    double range_inv = 1.0 / (P->max_dist - P->min_dist);
    double x = (2.0 * r - (P->min_dist + P->max_dist)) * range_inv;
    double dxdr = 2.0 * range_inv;
    chebyshev(N, x, T, dT);
    double s = (P->max_dist - r) * (P->max_dist - r);
    double ds = -2.0 * (P->max_dist - r);
    const double *c = P->radial_coeffs +
        ((size_t)(wrap_species(ti, P->Sp) * P->Sp + wrap_species(tk, P->Sp)) * R) * N;
    for (int m = 0; m < R; ++m) {
        double v = 0.0, dv = 0.0;
        for (int n = 0; n < N; ++n) { v += c[m * N + n] * T[n]; dv += c[m * N + n] * dT[n]; }
        f[m] = P->scaling * s * v;
        df[m] = P->scaling * (ds * v + s * dv * dxdr);
    }
    if (Tn) for (int n = 0; n < N; ++n) Tn[n] = P->scaling * s * T[n];
    return 1;
}

/* ------------------------------------------------------------------ */
/* Compact (MLIP alpha-table) form, conversion.py:39-102.              */
/* ------------------------------------------------------------------ */

static int max_pow_of(const orc_mtp *P)
{
    int mp = 0;
    for (int b = 0; b < P->B; ++b)
        for (int q = 1; q < 4; ++q) if (P->basic[4 * b + q] > mp) mp = P->basic[4 * b + q];
    return mp;
}

/* Basis values from unit vectors and given radial values rb[mu][k]
 * (conversion.py:39-84).  out[S].  Used for the coefficient remap. */
int orc_compact_basis_rb(const orc_mtp *P, int nn, const double *u, const double *rb, int rb_stride,
                         double *out)
{
    double *m = (double *)calloc((size_t)P->amc, sizeof(double));
    if (!m) return -1;
    int mp = max_pow_of(P);
    for (int b = 0; b < P->B; ++b) {
        int mu = P->basic[4 * b], a = P->basic[4 * b + 1], bb = P->basic[4 * b + 2], cc = P->basic[4 * b + 3];
        double acc = 0.0;
        for (int k = 0; k < nn; ++k) {
            double px = 1, py = 1, pz = 1;
            for (int q = 0; q < a; ++q) px *= u[3 * k];
            for (int q = 0; q < bb; ++q) py *= u[3 * k + 1];
            for (int q = 0; q < cc; ++q) pz *= u[3 * k + 2];
            acc += rb[mu * rb_stride + k] * (px * py * pz);
        }
        m[b] = acc;
    }
    (void)mp;
    for (int e = 0; e < P->T; ++e) {
        const int *t = P->times + 4 * e;
        m[t[3]] += t[2] * m[t[0]] * m[t[1]];
    }
    for (int s = 0; s < P->S; ++s) out[s] = m[P->mapping[s]];
    free(m);
    return 0;
}

/* One atom: energy, and optionally g[K][3], basis[S] and the radial
 * coefficient gradient dc[R][N] of this atom's species pair rows
 * (accumulated into dc_full[Sp][Sp][R][N] when given). */
static void compact_atom(const orc_mtp *P, int K, const double *rij, int ti, const int *tk,
                         double *Ei, double *g, double *basis, double *dc_full, double *work, const double *xi)
{
    if (!xi) xi = P->moment_coeffs;
    int R = P->R, N = P->N, amc = P->amc;
    int mp = max_pow_of(P);
    double *m = work;                 /* amc */
    double *dm = m + amc;             /* amc */
    double *f = dm + amc;             /* K*R */
    double *df = f + (size_t)K * R;   /* K*R */
    double *u = df + (size_t)K * R;   /* K*3 */
    double *r = u + (size_t)K * 3;    /* K */
    double *Tn = r + K;               /* K*N */
    int *valid = (int *)(Tn + (size_t)K * N); /* K */
    double pw[3][32];

    for (int k = 0; k < K; ++k) {
        const double *v = rij + 3 * k;
        r[k] = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        valid[k] = radial(P, r[k], ti, tk[k], f + k * R, df + k * R, Tn + (size_t)k * N);
        if (valid[k]) for (int q = 0; q < 3; ++q) u[3 * k + q] = v[q] / r[k];
        else u[3 * k] = u[3 * k + 1] = u[3 * k + 2] = 0.0;
    }
    memset(m, 0, sizeof(double) * amc);
    for (int k = 0; k < K; ++k) {
        if (!valid[k]) continue;
        for (int q = 0; q < 3; ++q) { pw[q][0] = 1.0; for (int p = 1; p <= mp; ++p) pw[q][p] = pw[q][p - 1] * u[3 * k + q]; }
        for (int b = 0; b < P->B; ++b) {
            const int *bs = P->basic + 4 * b;
            m[b] += f[k * R + bs[0]] * (pw[0][bs[1]] * pw[1][bs[2]] * pw[2][bs[3]]);
        }
    }
    for (int e = 0; e < P->T; ++e) {
        const int *t = P->times + 4 * e;
        m[t[3]] += t[2] * m[t[0]] * m[t[1]];
    }
    double E = P->species_coeffs[wrap_species(ti, P->Sp)];
    for (int s = 0; s < P->S; ++s) {
        E += xi[s] * m[P->mapping[s]];
        if (basis) basis[s] = m[P->mapping[s]];
    }
    *Ei = E;
    if (!g && !dc_full) return;

    /* reverse pass over the contraction table, reverse file order */
    memset(dm, 0, sizeof(double) * amc);
    for (int s = 0; s < P->S; ++s) dm[P->mapping[s]] += xi[s];
    for (int e = P->T - 1; e >= 0; --e) {
        const int *t = P->times + 4 * e;
        double w = t[2] * dm[t[3]];
        dm[t[0]] += w * m[t[1]];
        dm[t[1]] += w * m[t[0]];
    }
    for (int k = 0; k < K; ++k) {
        double gx = 0, gy = 0, gz = 0;
        if (!valid[k]) { if (g) { g[3 * k] = g[3 * k + 1] = g[3 * k + 2] = 0.0; } continue; }
        double ux = u[3 * k], uy = u[3 * k + 1], uz = u[3 * k + 2];
        for (int q = 0; q < 3; ++q) { pw[q][0] = 1.0; for (int p = 1; p <= mp; ++p) pw[q][p] = pw[q][p - 1] * u[3 * k + q]; }
        double Amu[16] = {0}; /* sum_b dm[b] P_b, per mu */
        double sdf = 0.0;     /* sum_b dm[b] f'_mu P_b */
        double sfn = 0.0;     /* sum_b dm[b] f_mu nu P_b */
        double G[3] = {0, 0, 0}; /* sum_b dm[b] f_mu grad_u P_b */
        for (int b = 0; b < P->B; ++b) {
            const int *bs = P->basic + 4 * b;
            int mu = bs[0], a = bs[1], bb = bs[2], cc = bs[3];
            double d = dm[b];
            double Pb = pw[0][a] * pw[1][bb] * pw[2][cc];
            Amu[mu] += d * Pb;
            double fm = f[k * R + mu];
            sdf += d * df[k * R + mu] * Pb;
            sfn += d * fm * (double)(a + bb + cc) * Pb;
            if (a > 0) G[0] += d * fm * a * pw[0][a - 1] * pw[1][bb] * pw[2][cc];
            if (bb > 0) G[1] += d * fm * bb * pw[0][a] * pw[1][bb - 1] * pw[2][cc];
            if (cc > 0) G[2] += d * fm * cc * pw[0][a] * pw[1][bb] * pw[2][cc - 1];
        }
        double ir = 1.0 / r[k];
        gx = sdf * ux + (G[0] - sfn * ux) * ir;
        gy = sdf * uy + (G[1] - sfn * uy) * ir;
        gz = sdf * uz + (G[2] - sfn * uz) * ir;
        if (g) { g[3 * k] = gx; g[3 * k + 1] = gy; g[3 * k + 2] = gz; }
        if (dc_full) {
            double *dc = dc_full + ((size_t)(wrap_species(ti, P->Sp) * P->Sp + wrap_species(tk[k], P->Sp)) * R) * N;
            for (int mu = 0; mu < R; ++mu)
                for (int n = 0; n < N; ++n) dc[mu * N + n] += Amu[mu] * Tn[(size_t)k * N + n];
        }
    }
}

static size_t compact_work_doubles(const orc_mtp *P, int K)
{
    return (size_t)2 * P->amc + (size_t)K * (2 * P->R + 4 + P->N) + (size_t)K + 16;
}

/* Per-atom energies Ei[A] and per-slot gradients g[A][K][3] (may be NULL)
 * for A padded rows of K slots (jax_backend.py:20-57 semantics). */
int orc_compact_atoms(const orc_mtp *P, int A, int K, const int *itypes, const double *rij,
                      const int *jtypes, double *Ei, double *g, int nthreads)
{
    if (P->R > 16 || P->N > 64) return -2;
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *work = (double *)malloc(sizeof(double) * compact_work_doubles(P, K));
        if (!work) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            err = -1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
            for (int i = 0; i < A; ++i)
                compact_atom(P, K, rij + (size_t)i * K * 3, itypes[i], jtypes + (size_t)i * K, Ei + i,
                             g ? g + (size_t)i * K * 3 : NULL, NULL, NULL, work, NULL);
            free(work);
        }
    }
    return err;
}

/* Per-atom basis vectors (B_mlip, file order) and the gradient of the
 * summed energy of rows [0, n_rows) w.r.t. radial coefficients.  Single
 * threaded; for parameter-gradient checks on small configurations. */
int orc_compact_basis_and_dc(const orc_mtp *P, int A, int K, const int *itypes, const double *rij,
                             const int *jtypes, double *Ei, double *basis, double *dc_full)
{
    double *work = (double *)malloc(sizeof(double) * compact_work_doubles(P, K));
    if (!work) return -1;
    for (int i = 0; i < A; ++i)
        compact_atom(P, K, rij + (size_t)i * K * 3, itypes[i], jtypes + (size_t)i * K, Ei + i, NULL,
                     basis ? basis + (size_t)i * P->S : NULL, dc_full, work, NULL);
    free(work);
    return 0;
}

/* Per-row coefficient vectors xi_row = xi + i * xi_stride in place of
 * moment_coeffs: Ei = species_coeff + xi_row . b_i, g = its slot gradient,
 * basis (may be NULL) = b_i.  Basis Jacobians and the distilled head's
 * forces are built from this (the reference has no counterpart). */
int orc_compact_atoms_xi(const orc_mtp *P, int A, int K, const int *itypes, const double *rij, const int *jtypes,
                         const double *xi, int xi_stride, double *Ei, double *g, double *basis, int nthreads)
{
    if (P->R > 16 || P->N > 64) return -2;
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *work = (double *)malloc(sizeof(double) * compact_work_doubles(P, K));
        if (!work) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            err = -1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
            for (int i = 0; i < A; ++i)
                compact_atom(P, K, rij + (size_t)i * K * 3, itypes[i], jtypes + (size_t)i * K, Ei + i,
                             g ? g + (size_t)i * K * 3 : NULL, basis ? basis + (size_t)i * P->S : NULL, NULL, work,
                             xi + (size_t)i * xi_stride);
            free(work);
        }
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* Full-tensor tree form, jax_backend.py:162-247.                      */
/* ------------------------------------------------------------------ */

typedef struct {
    int nnodes, nscalar, max_nu;
    int *kind, *mu, *nu, *rank, *left, *right;
    size_t *off;        /* offset of node tensor in value buffer */
    size_t total;       /* total tensor elements */
    /* contraction offset tables */
    int **offA_free, **offA_con, **offB_free, **offB_con;
    int *nfa, *nfb, *ncon;
    int *scalar_node;   /* [nscalar], -1 -> 0.0 (jax_backend.py:215) */
} orc_tree;

static int ipow3(int p) { int v = 1; while (p-- > 0) v *= 3; return v; }

/* Offsets of a rank-`rank` row-major 3^rank tensor for all index tuples of
 * the axes listed in `axes` (others zero), enumerated row-major over those
 * axes in the listed order. */
static int *axis_offsets(int rank, const int *axes, int nax)
{
    int cnt = ipow3(nax);
    int *o = (int *)malloc(sizeof(int) * (size_t)cnt);
    int idx[16] = {0};
    for (int c = 0; c < cnt; ++c) {
        int rem = c;
        for (int t = nax - 1; t >= 0; --t) { idx[t] = rem % 3; rem /= 3; }
        int off = 0;
        for (int t = 0; t < nax; ++t) off += idx[t] * ipow3(rank - 1 - axes[t]);
        o[c] = off;
    }
    return o;
}

/* Build the tree from node arrays in execution order.
 * kind[i]: 0 basic (mu, nu), 1 contraction (left, right, naxes, laxes, raxes)
 * laxes/raxes: [nnodes][16]. */
orc_tree *orc_tree_build(int nnodes, const int *kind, const int *mu, const int *nu, const int *left,
                         const int *right, const int *naxes, const int *laxes, const int *raxes,
                         int nscalar, const int *scalar_node)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->nnodes = nnodes; t->nscalar = nscalar;
    t->kind = (int *)malloc(sizeof(int) * nnodes); t->mu = (int *)malloc(sizeof(int) * nnodes);
    t->nu = (int *)malloc(sizeof(int) * nnodes); t->rank = (int *)malloc(sizeof(int) * nnodes);
    t->left = (int *)malloc(sizeof(int) * nnodes); t->right = (int *)malloc(sizeof(int) * nnodes);
    t->off = (size_t *)malloc(sizeof(size_t) * nnodes);
    t->offA_free = (int **)calloc(nnodes, sizeof(int *)); t->offA_con = (int **)calloc(nnodes, sizeof(int *));
    t->offB_free = (int **)calloc(nnodes, sizeof(int *)); t->offB_con = (int **)calloc(nnodes, sizeof(int *));
    t->nfa = (int *)calloc(nnodes, sizeof(int)); t->nfb = (int *)calloc(nnodes, sizeof(int));
    t->ncon = (int *)calloc(nnodes, sizeof(int));
    t->scalar_node = (int *)malloc(sizeof(int) * (nscalar > 0 ? nscalar : 1));
    memcpy(t->scalar_node, scalar_node, sizeof(int) * nscalar);
    size_t tot = 0;
    t->max_nu = 0;
    for (int i = 0; i < nnodes; ++i) {
        t->kind[i] = kind[i]; t->mu[i] = mu[i]; t->nu[i] = nu[i];
        t->left[i] = left[i]; t->right[i] = right[i];
        if (kind[i] == 0) {
            t->rank[i] = nu[i];
            if (nu[i] > t->max_nu) t->max_nu = nu[i];
        } else {
            int ra = t->rank[left[i]], rb = t->rank[right[i]], n = naxes[i];
            t->rank[i] = ra + rb - 2 * n;
            const int *la = laxes + 16 * i, *rx = raxes + 16 * i;
            int fa[16], fb[16], nf_a = 0, nf_b = 0;
            for (int q = 0; q < ra; ++q) { int used = 0; for (int s = 0; s < n; ++s) if (la[s] == q) used = 1; if (!used) fa[nf_a++] = q; }
            for (int q = 0; q < rb; ++q) { int used = 0; for (int s = 0; s < n; ++s) if (rx[s] == q) used = 1; if (!used) fb[nf_b++] = q; }
            t->offA_free[i] = axis_offsets(ra, fa, nf_a); t->offA_con[i] = axis_offsets(ra, la, n);
            t->offB_free[i] = axis_offsets(rb, fb, nf_b); t->offB_con[i] = axis_offsets(rb, rx, n);
            t->nfa[i] = ipow3(nf_a); t->nfb[i] = ipow3(nf_b); t->ncon[i] = ipow3(n);
        }
        t->off[i] = tot;
        tot += (size_t)ipow3(t->rank[i]);
    }
    t->total = tot;
    return t;
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    for (int i = 0; i < t->nnodes; ++i) {
        free(t->offA_free[i]); free(t->offA_con[i]); free(t->offB_free[i]); free(t->offB_con[i]);
    }
    free(t->offA_free); free(t->offA_con); free(t->offB_free); free(t->offB_con);
    free(t->nfa); free(t->nfb); free(t->ncon);
    free(t->kind); free(t->mu); free(t->nu); free(t->rank); free(t->left); free(t->right);
    free(t->off); free(t->scalar_node); free(t);
}

long long orc_tree_total(const orc_tree *t) { return (long long)t->total; }
int orc_tree_rank(const orc_tree *t, int node) { return t->rank[node]; }

/* Contractions (tensordot) forward, in node order. */
static void tree_contract_fwd(const orc_tree *t, double *val)
{
    for (int i = 0; i < t->nnodes; ++i) {
        if (t->kind[i] != 1) continue;
        const double *A = val + t->off[t->left[i]], *Bv = val + t->off[t->right[i]];
        double *C = val + t->off[i];
        int nfa = t->nfa[i], nfb = t->nfb[i], nc = t->ncon[i];
        const int *oaf = t->offA_free[i], *oac = t->offA_con[i], *obf = t->offB_free[i], *obc = t->offB_con[i];
        for (int x = 0; x < nfa; ++x)
            for (int y = 0; y < nfb; ++y) {
                double s = 0.0;
                for (int c = 0; c < nc; ++c) s += A[oaf[x] + oac[c]] * Bv[obf[y] + obc[c]];
                C[(size_t)x * nfb + y] = s;
            }
    }
}

static void tree_contract_rev(const orc_tree *t, const double *val, double *adj)
{
    for (int i = t->nnodes - 1; i >= 0; --i) {
        if (t->kind[i] != 1) continue;
        const double *A = val + t->off[t->left[i]], *Bv = val + t->off[t->right[i]];
        double *dA = adj + t->off[t->left[i]], *dB = adj + t->off[t->right[i]];
        const double *dC = adj + t->off[i];
        int nfa = t->nfa[i], nfb = t->nfb[i], nc = t->ncon[i];
        const int *oaf = t->offA_free[i], *oac = t->offA_con[i], *obf = t->offB_free[i], *obc = t->offB_con[i];
        for (int x = 0; x < nfa; ++x)
            for (int y = 0; y < nfb; ++y) {
                double d = dC[(size_t)x * nfb + y];
                if (d == 0.0) continue;
                for (int c = 0; c < nc; ++c) {
                    dA[oaf[x] + oac[c]] += d * Bv[obf[y] + obc[c]];
                    dB[obf[y] + obc[c]] += d * A[oaf[x] + oac[c]];
                }
            }
    }
}

/* u^{(x)p} for p = 0..max_nu, stored consecutively (make_tensor, utils.py:16-27). */
static void outer_powers(const double *u, int max_nu, double *U, size_t *Uoff)
{
    size_t o = 0;
    Uoff[0] = 0; U[0] = 1.0; o = 1;
    for (int p = 1; p <= max_nu; ++p) {
        Uoff[p] = o;
        int prev = ipow3(p - 1);
        const double *Up = U + Uoff[p - 1];
        for (int J = 0; J < prev; ++J)
            for (int q = 0; q < 3; ++q) U[o + (size_t)J * 3 + q] = Up[J] * u[q];
        o += (size_t)prev * 3;
    }
}

/* Basis values from unit vectors and given rb[mu][k] (conversion.py:171-198). */
int orc_tree_basis_rb(const orc_tree *t, int nn, const double *u, const double *rb, int rb_stride, double *out)
{
    double *val = (double *)calloc(t->total, sizeof(double));
    size_t usz = 0; for (int p = 0; p <= t->max_nu; ++p) usz += (size_t)ipow3(p);
    double *U = (double *)malloc(sizeof(double) * usz);
    size_t Uoff[32];
    for (int k = 0; k < nn; ++k) {
        outer_powers(u + 3 * k, t->max_nu, U, Uoff);
        for (int i = 0; i < t->nnodes; ++i) {
            if (t->kind[i] != 0) continue;
            double w = rb[t->mu[i] * rb_stride + k];
            int sz = ipow3(t->nu[i]);
            const double *Un = U + Uoff[t->nu[i]];
            double *M = val + t->off[i];
            for (int e = 0; e < sz; ++e) M[e] += w * Un[e];
        }
    }
    tree_contract_fwd(t, val);
    for (int s = 0; s < t->nscalar; ++s) out[s] = t->scalar_node[s] >= 0 ? val[t->off[t->scalar_node[s]]] : 0.0;
    free(val); free(U);
    return 0;
}

static void tree_atom(const orc_mtp *P, const orc_tree *t, const double *xi_tree, int K, const double *rij,
                      int ti, const int *tk, double *Ei, double *g, double *val, double *adj, double *U,
                      double *D, double *Rb)
{
    int R = P->R;
    size_t Uoff[32];
    double f[16], df[16];
    memset(val, 0, sizeof(double) * t->total);
    for (int k = 0; k < K; ++k) {
        const double *v = rij + 3 * k;
        double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!radial(P, r, ti, tk[k], f, df, NULL)) continue;
        double u[3] = {v[0] / r, v[1] / r, v[2] / r};
        outer_powers(u, t->max_nu, U, Uoff);
        for (int i = 0; i < t->nnodes; ++i) {
            if (t->kind[i] != 0) continue;
            double w = f[t->mu[i]];
            int sz = ipow3(t->nu[i]);
            const double *Un = U + Uoff[t->nu[i]];
            double *M = val + t->off[i];
            for (int e = 0; e < sz; ++e) M[e] += w * Un[e];
        }
    }
    tree_contract_fwd(t, val);
    double E = P->species_coeffs[wrap_species(ti, P->Sp)];
    for (int s = 0; s < t->nscalar; ++s)
        if (t->scalar_node[s] >= 0) E += xi_tree[s] * val[t->off[t->scalar_node[s]]];
    *Ei = E;
    if (!g) return;
    memset(adj, 0, sizeof(double) * t->total);
    for (int s = 0; s < t->nscalar; ++s)
        if (t->scalar_node[s] >= 0) adj[t->off[t->scalar_node[s]]] += xi_tree[s];
    tree_contract_rev(t, val, adj);
    for (int k = 0; k < K; ++k) {
        const double *v = rij + 3 * k;
        double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!radial(P, r, ti, tk[k], f, df, NULL)) { g[3 * k] = g[3 * k + 1] = g[3 * k + 2] = 0.0; continue; }
        double u[3] = {v[0] / r, v[1] / r, v[2] / r};
        outer_powers(u, t->max_nu, U, Uoff);
        double dEdf[16] = {0};
        double du[3] = {0, 0, 0};
        for (int nuv = 0; nuv <= t->max_nu; ++nuv) {
            int sz = ipow3(nuv), any = 0;
            const double *Un = U + Uoff[nuv];
            for (int e = 0; e < sz; ++e) D[e] = 0.0;
            for (int i = 0; i < t->nnodes; ++i) {
                if (t->kind[i] != 0 || t->nu[i] != nuv) continue;
                any = 1;
                const double *dM = adj + t->off[i];
                double a = 0.0, w = f[t->mu[i]];
                for (int e = 0; e < sz; ++e) { a += dM[e] * Un[e]; D[e] += w * dM[e]; }
                dEdf[t->mu[i]] += a;
            }
            if (!any || nuv == 0) continue;
            /* grad_u (D . u^{(x)nu}) by the contraction chain: R_nu = D,
             * R_p[J] = sum_i R_{p+1}[J,i] u_i, grad_i = sum_p U_p . R_{p+1}[:, i] */
            double *Rcur = Rb;
            memcpy(Rcur, D, sizeof(double) * sz);
            int cur = sz; /* size of R_{p+1} */
            for (int p = nuv - 1; p >= 0; --p) {
                int np = cur / 3;
                const double *Up = U + Uoff[p];
                for (int J = 0; J < np; ++J)
                    for (int q = 0; q < 3; ++q) du[q] += Up[J] * Rcur[(size_t)J * 3 + q];
                double *Rn = Rcur + cur;
                for (int J = 0; J < np; ++J)
                    Rn[J] = Rcur[(size_t)J * 3] * u[0] + Rcur[(size_t)J * 3 + 1] * u[1] + Rcur[(size_t)J * 3 + 2] * u[2];
                Rcur = Rn; cur = np;
            }
        }
        double sdf = 0.0;
        for (int m = 0; m < R; ++m) sdf += dEdf[m] * df[m];
        double ud = u[0] * du[0] + u[1] * du[1] + u[2] * du[2];
        for (int q = 0; q < 3; ++q) g[3 * k + q] = sdf * u[q] + (du[q] - ud * u[q]) / r;
    }
}

int orc_tree_atoms(const orc_mtp *P, const orc_tree *t, const double *xi_tree, int A, int K, const int *itypes,
                   const double *rij, const int *jtypes, double *Ei, double *g, int nthreads)
{
    if (P->R > 16) return -2;
    size_t usz = 0; for (int p = 0; p <= t->max_nu; ++p) usz += (size_t)ipow3(p);
    size_t dsz = (size_t)ipow3(t->max_nu);
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *val = (double *)malloc(sizeof(double) * t->total);
        double *adj = (double *)malloc(sizeof(double) * t->total);
        double *U = (double *)malloc(sizeof(double) * usz);
        double *D = (double *)malloc(sizeof(double) * dsz);
        double *Rb = (double *)malloc(sizeof(double) * (2 * dsz + 4));
        if (!val || !adj || !U || !D || !Rb) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            err = -1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int i = 0; i < A; ++i)
                tree_atom(P, t, xi_tree, K, rij + (size_t)i * K * 3, itypes[i], jtypes + (size_t)i * K, Ei + i,
                          g ? g + (size_t)i * K * 3 : NULL, val, adj, U, D, Rb);
        }
        free(val); free(adj); free(U); free(D); free(Rb);
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* Force scatter, stress, masked energy.                               */
/* ------------------------------------------------------------------ */

/* jax_backend.py:263-303: F_a = [a<n] sum_k g_ak - sum_{i<n, k: j_ik=a<n} g_ik */
void orc_accumulate_forces(int A, int K, const double *g, const int *js, int natoms, double *F)
{
    memset(F, 0, sizeof(double) * (size_t)A * 3);
    for (int i = 0; i < A && i < natoms; ++i)
        for (int k = 0; k < K; ++k) {
            const double *gk = g + ((size_t)i * K + k) * 3;
            F[3 * i] += gk[0]; F[3 * i + 1] += gk[1]; F[3 * i + 2] += gk[2];
            int j = js[(size_t)i * K + k];
            if (j < natoms && j >= 0 && j < A) {
                F[3 * j] -= gk[0]; F[3 * j + 1] -= gk[1]; F[3 * j + 2] -= gk[2];
            }
        }
}

/* jax_backend.py:355-370: W = sum r (x) g over all rows; sigma = (W+W^T)/(2V);
 * Voigt [xx, yy, zz, yz, xz, xy]; NaN unless cell_rank == 3. Also W[9]. */
void orc_stress(int A, int K, const double *rij, const double *g, int cell_rank, double volume,
                double *stress6, double *W)
{
    double w[9] = {0};
    for (size_t s = 0; s < (size_t)A * K; ++s)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) w[3 * a + b] += rij[3 * s + a] * g[3 * s + b];
    if (W) memcpy(W, w, sizeof w);
    static const int idx[6] = {0, 4, 8, 5, 2, 1};
    for (int q = 0; q < 6; ++q) {
        int a = idx[q] / 3, b = idx[q] % 3;
        stress6[q] = cell_rank == 3 ? (w[3 * a + b] + w[3 * b + a]) * (0.5 / volume) : NAN;
    }
}

/* jax_backend.py:431-456 */
double orc_energy(int A, const double *Ei, int natoms_energy)
{
    double s = 0.0;
    for (int i = 0; i < A && i < natoms_energy; ++i) s += Ei[i];
    return s;
}

/* ------------------------------------------------------------------ */
/* Neighbor packing of the pair style, pair_jax_mtp.cpp compute().     */
/* ------------------------------------------------------------------ */

/* LAMMPS Domain::minimum_image for an orthogonal periodic box. */
static void min_image(double *d, const double *prd, const int *periodic)
{
    for (int q = 0; q < 3; ++q) {
        if (!periodic[q]) continue;
        if (fabs(d[q]) > 0.5 * prd[q]) {
            if (d[q] < 0.0) d[q] += prd[q];
            else d[q] -= prd[q];
        }
    }
}

/* Restates pair_jax_mtp.cpp:96-173 (compute() body) for a full neighbor
 * list: produces per-local-atom rows of 3*max_nbr doubles (pos), max_nbr
 * ints (nbr, ntype), counts[nlocal], atom_types[max_atoms] (type-1, -1
 * beyond nlocal) and returns nneigh_total.  Rows are written for the atoms
 * listed in ilist only, like the reference. */
long long orc_pack(int nlocal, int nall, const double *x, const int *type, const int *tag, int inum,
                   const int *ilist, const int *numneigh, const int *const *firstneigh, double cutoff,
                   const double *prd, const int *periodic, int max_atoms, int max_nbr, double *pos, int *nbr,
                   int *ntype, int *counts, int *atom_types)
{
    (void)nall;
    const double cutsq = cutoff * cutoff;
    for (int i = 0; i < max_atoms; ++i) { atom_types[i] = -1; counts[i] = 0; }
    for (int i = 0; i < nlocal; ++i) atom_types[i] = type[i] - 1;
    int max_tag = 0;
    for (int i = 0; i < nlocal; ++i) if (tag[i] > max_tag) max_tag = tag[i];
    int *t2d = (int *)malloc(sizeof(int) * (size_t)(max_tag + 1));
    for (int q = 0; q <= max_tag; ++q) t2d[q] = -1;
    for (int i = 0; i < nlocal; ++i) if (tag[i] >= 0 && tag[i] <= max_tag) t2d[tag[i]] = i;
    long long total = 0;
    for (int ii = 0; ii < inum; ++ii) {
        int i = ilist[ii];
        int cnt = 0;
        double *prow = pos + (size_t)i * 3 * max_nbr;
        int *nrow = nbr + (size_t)i * max_nbr, *trow = ntype + (size_t)i * max_nbr;
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            int j = firstneigh[i][jj] & 0x1FFFFFFF; /* NEIGHMASK */
            if (j == i) continue;
            double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
            min_image(d, prd, periodic);
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (rsq < cutsq && cnt < max_nbr) {
                prow[3 * cnt] = d[0]; prow[3 * cnt + 1] = d[1]; prow[3 * cnt + 2] = d[2];
                int dense = nlocal;
                int jt = tag[j];
                if (jt >= 0 && jt <= max_tag && t2d[jt] != -1) dense = t2d[jt];
                nrow[cnt] = dense;
                trow[cnt] = dense != nlocal ? type[j] - 1 : -1;
                ++cnt;
            }
        }
        counts[i] = cnt;
        total += cnt;
        for (int q = cnt; q < max_nbr; ++q) {
            prow[3 * q] = prow[3 * q + 1] = prow[3 * q + 2] = 20.0;
            nrow[q] = nlocal; trow[q] = -1;
        }
    }
    free(t2d);
    return total;
}

/* Restates zero_overhead.cpp:61-117 (update_atom_data_zero_copy) without
 * the float cast of r: every slot of the padded [max_atoms][max_nbr] arrays
 * gets the sentinels (j = natoms, type = -1, r = 2000, itype = -1), then the
 * counts[i] valid prefix of each row i < natoms is copied. */
void orc_stage(int max_atoms, int max_nbr, int natoms, const double *pos, const int *atom_types, const int *nbr,
               const int *counts, const int *ntype, int *itypes_out, int *js_out, double *rij_out,
               int *jtypes_out)
{
    size_t tot = (size_t)max_atoms * max_nbr;
    for (size_t q = 0; q < tot; ++q) { js_out[q] = natoms; jtypes_out[q] = -1; }
    for (size_t q = 0; q < tot * 3; ++q) rij_out[q] = 2000.0;
    for (int i = 0; i < max_atoms; ++i) itypes_out[i] = -1;
    for (int i = 0; i < natoms; ++i) itypes_out[i] = atom_types[i];
    for (int i = 0; i < natoms; ++i) {
        size_t base = (size_t)i * max_nbr;
        for (int j = 0; j < counts[i]; ++j) {
            js_out[base + j] = nbr[base + j];
            jtypes_out[base + j] = ntype[base + j];
            for (int q = 0; q < 3; ++q) rij_out[(base + j) * 3 + q] = pos[(base + j) * 3 + q];
        }
    }
}
