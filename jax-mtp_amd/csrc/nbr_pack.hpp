// Neighbor packing of the pair style, shared by the LAMMPS pair style
// (jax-mtp_amd/lammps/pair_jax_mtp.cpp) and the C ABI helper
// mtp_pack_full_list.
//
// Semantics follow PairJaxMTP::compute (pair_jax_mtp.cpp:96-173):
// per local atom i of the full list, skip j == i, take del = x[j] - x[i]
// through the box's minimum image, keep it if r^2 < cutoff^2 while fewer
// than max_nbr are kept (further in-cutoff neighbors are dropped and
// counted), map j to a local index through its tag (ghosts fold onto their
// owner; unmapped -> nlocal with type -1), store type[j] - 1, and pad the
// rest of the row with (20, 20, 20) / nlocal / -1.
#pragma once
#include <cstdint>
#include <vector>

namespace mtp {

struct PackOut {
    double* pos;     // [max_atoms][3*max_nbr]
    int* nbr;        // [max_atoms][max_nbr]
    int* ntype;      // [max_atoms][max_nbr]
    int* counts;     // [max_atoms]
    int* atom_types; // [max_atoms]
};

// MinImage: callable (double& dx, double& dy, double& dz).  TagT is LAMMPS's
// tagint (int, or int64_t in LAMMPS_BIGBIG builds).
template <typename MinImage, typename TagT>
long long pack_full_list(int nlocal, const double* const* x, const int* type, const TagT* tag, int inum,
                         const int* ilist, const int* numneigh, const int* const* firstneigh, double cutoff,
                         int max_atoms, int max_nbr, MinImage&& min_image, const PackOut& o,
                         std::vector<int>& tag_to_dense, long long* n_truncated)
{
    constexpr int kNeighMask = 0x1FFFFFFF;  // LAMMPS NEIGHMASK
    const double cutsq = cutoff * cutoff;
    for (int i = 0; i < max_atoms; ++i) { o.atom_types[i] = -1; o.counts[i] = 0; }
    for (int i = 0; i < nlocal; ++i) o.atom_types[i] = type[i] - 1;
    long long max_tag = 0;
    for (int i = 0; i < nlocal; ++i) max_tag = (long long)tag[i] > max_tag ? (long long)tag[i] : max_tag;
    tag_to_dense.assign((size_t)max_tag + 1, -1);
    for (int i = 0; i < nlocal; ++i)
        if (tag[i] >= 0 && tag[i] <= max_tag) tag_to_dense[tag[i]] = i;
    long long total = 0, dropped = 0;
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        const double xi = x[i][0], yi = x[i][1], zi = x[i][2];
        double* prow = o.pos + (size_t)i * 3 * max_nbr;
        int* nrow = o.nbr + (size_t)i * max_nbr;
        int* trow = o.ntype + (size_t)i * max_nbr;
        int cnt = 0;
        const int* jlist = firstneigh[i];
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            const int j = jlist[jj] & kNeighMask;
            if (j == i) continue;
            double dx = x[j][0] - xi, dy = x[j][1] - yi, dz = x[j][2] - zi;
            min_image(dx, dy, dz);
            const double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < cutsq) {
                if (cnt >= max_nbr) { ++dropped; continue; }
                prow[3 * cnt] = dx; prow[3 * cnt + 1] = dy; prow[3 * cnt + 2] = dz;
                int dense = nlocal;
                const long long jt = (long long)tag[j];
                if (jt >= 0 && jt <= max_tag && tag_to_dense[jt] != -1) dense = tag_to_dense[jt];
                nrow[cnt] = dense;
                trow[cnt] = dense != nlocal ? type[j] - 1 : -1;
                ++cnt;
            }
        }
        o.counts[i] = cnt;
        total += cnt;
        for (int q = cnt; q < max_nbr; ++q) {
            prow[3 * q] = prow[3 * q + 1] = prow[3 * q + 2] = 20.0;
            nrow[q] = nlocal;
            trow[q] = -1;
        }
    }
    if (n_truncated) *n_truncated = dropped;
    return total;
}

// LAMMPS Domain::minimum_image for an orthogonal box.
struct OrthoMinImage {
    double prd[3];
    int periodic[3];
    void operator()(double& dx, double& dy, double& dz) const
    {
        double* d[3] = {&dx, &dy, &dz};
        for (int q = 0; q < 3; ++q) {
            if (!periodic[q]) continue;
            double& v = *d[q];
            if ((v < 0 ? -v : v) > 0.5 * prd[q]) v += v < 0.0 ? prd[q] : -prd[q];
        }
    }
};

}  // namespace mtp
