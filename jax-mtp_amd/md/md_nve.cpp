// Standalone velocity-Verlet NVE driver (LAMMPS metal units) over the
// pair-style host path, for running an MTP where LAMMPS is absent.
//
// Each step does what LAMMPS + pair_style jax/mtp would do: local atoms in a
// periodic orthogonal box, ghost copies of atoms near the faces carrying the
// owners' tags, a full neighbor list with a skin (rebuilt every step), the
// pair style's packing (mtp::pack_full_list, pair_jax_mtp.cpp:96-173), then
// mtp_compute() on the GPU, then the fix-nve half-kicks and drift.
//
// Usage:
//   md_nve <file.mtp> <lattice bcc|fcc> <a> <nx> <ny> <nz> <natoms|0> <steps> <dt_ps>
//          <T_K> <sigma> <seed> <out_prefix> [dump_every] [mass0 mass1 ...]
// Writes <out_prefix>.thermo (step pe ke etot) and, every dump_every steps,
// <out_prefix>.dumpNNNNNN (binary: natoms, box[3], types, positions, forces).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mtp_hip.h"
#include "nbr_pack.hpp"

namespace {

constexpr double kFtm2v = 1.0 / 1.0364269e-4;  // metal: (eV/A) / (g/mol) -> A/ps^2
constexpr double kMvv2e = 1.0364269e-4;        // metal: (g/mol) A^2/ps^2 -> eV
constexpr double kBoltz = 8.617343e-5;         // eV/K

struct System {
    int nlocal = 0;
    double L[3] = {0, 0, 0};
    std::vector<double> x;   // [nall][3] (locals first, then ghosts)
    std::vector<int> type;   // 1-based, LAMMPS style
    std::vector<int> tag;
    std::vector<double> v, f, mass;
    int nall() const { return (int)type.size(); }
};

void build_lattice(System& s, const char* kind, double a, int nx, int ny, int nz, int natoms, double sigma,
                   std::mt19937_64& rng)
{
    std::vector<std::array<double, 3>> base;
    if (std::strcmp(kind, "bcc") == 0) base = {{0, 0, 0}, {0.5, 0.5, 0.5}};
    else base = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    std::normal_distribution<double> nd(0.0, sigma);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                for (auto& b : base) {
                    if (natoms > 0 && s.nlocal >= natoms) continue;
                    s.x.push_back((i + b[0]) * a + nd(rng));
                    s.x.push_back((j + b[1]) * a + nd(rng));
                    s.x.push_back((k + b[2]) * a + nd(rng));
                    ++s.nlocal;
                }
    s.L[0] = nx * a; s.L[1] = ny * a; s.L[2] = nz * a;
}

void wrap(System& s)
{
    for (int i = 0; i < s.nlocal; ++i)
        for (int q = 0; q < 3; ++q) {
            double& c = s.x[3 * i + q];
            c -= s.L[q] * std::floor(c / s.L[q]);
        }
}

// ghosts: periodic copies of locals within rg of the box, same tag/type.
void make_ghosts(System& s, double rg)
{
    s.x.resize(3 * (size_t)s.nlocal);
    s.type.resize(s.nlocal);
    s.tag.resize(s.nlocal);
    for (int i = 0; i < s.nlocal; ++i)
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int c = -1; c <= 1; ++c) {
                    if (!a && !b && !c) continue;
                    const double p[3] = {s.x[3 * i] + a * s.L[0], s.x[3 * i + 1] + b * s.L[1],
                                         s.x[3 * i + 2] + c * s.L[2]};
                    bool in = true;
                    for (int q = 0; q < 3; ++q) in = in && p[q] >= -rg && p[q] < s.L[q] + rg;
                    if (!in) continue;
                    s.x.insert(s.x.end(), p, p + 3);
                    s.type.push_back(s.type[i]);
                    s.tag.push_back(s.tag[i]);
                }
}

// full neighbor list of locals over locals + ghosts within rn (binned)
void neighbor_list(const System& s, double rn, std::vector<std::vector<int>>& nl)
{
    const int nall = s.nall();
    int nb[3];
    double lo[3], bs[3];
    for (int q = 0; q < 3; ++q) {
        lo[q] = -rn;
        nb[q] = std::max(1, (int)std::floor((s.L[q] + 2 * rn) / rn));
        bs[q] = (s.L[q] + 2 * rn) / nb[q];
    }
    auto bin_of = [&](const double* p, int* b) {
        for (int q = 0; q < 3; ++q) b[q] = std::min(nb[q] - 1, std::max(0, (int)((p[q] - lo[q]) / bs[q])));
    };
    std::vector<std::vector<int>> bins((size_t)nb[0] * nb[1] * nb[2]);
    for (int j = 0; j < nall; ++j) {
        int b[3];
        bin_of(&s.x[3 * j], b);
        bins[((size_t)b[0] * nb[1] + b[1]) * nb[2] + b[2]].push_back(j);
    }
    nl.assign(s.nlocal, {});
    const double rn2 = rn * rn;
    for (int i = 0; i < s.nlocal; ++i) {
        int b[3];
        bin_of(&s.x[3 * i], b);
        for (int a = b[0] - 1; a <= b[0] + 1; ++a)
            for (int c = b[1] - 1; c <= b[1] + 1; ++c)
                for (int d = b[2] - 1; d <= b[2] + 1; ++d) {
                    if (a < 0 || c < 0 || d < 0 || a >= nb[0] || c >= nb[1] || d >= nb[2]) continue;
                    for (int j : bins[((size_t)a * nb[1] + c) * nb[2] + d]) {
                        if (j == i) continue;
                        double r2 = 0;
                        for (int q = 0; q < 3; ++q) { const double t = s.x[3 * j + q] - s.x[3 * i + q]; r2 += t * t; }
                        if (r2 < rn2) nl[i].push_back(j);
                    }
                }
    }
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 14) {
        std::fprintf(stderr, "usage: md_nve <file.mtp> <bcc|fcc> <a> <nx> <ny> <nz> <natoms|0> <steps> <dt_ps> <T_K> "
                             "<sigma> <seed> <out_prefix> [dump_every] [mass per type ...]\n");
        return 2;
    }
    const char* mtp_path = argv[1];
    const char* kind = argv[2];
    const double a = std::atof(argv[3]);
    const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]), nz = std::atoi(argv[6]);
    const int natoms_req = std::atoi(argv[7]);
    const int steps = std::atoi(argv[8]);
    const double dt = std::atof(argv[9]);
    const double T0 = std::atof(argv[10]);
    const double sigma = std::atof(argv[11]);
    const unsigned long long seed = std::strtoull(argv[12], nullptr, 10);
    const std::string out = argv[13];
    const int dump_every = argc > 14 ? std::atoi(argv[14]) : 0;
    std::vector<double> type_mass;
    for (int q = 15; q < argc; ++q) type_mass.push_back(std::atof(argv[q]));
    if (type_mass.empty()) type_mass = {58.6934, 26.9815};

    std::mt19937_64 rng(seed);
    System s;
    build_lattice(s, kind, a, nx, ny, nz, natoms_req, sigma, rng);
    const int n = s.nlocal;
    const int ntypes = (int)type_mass.size();
    // species: equal shares by seeded shuffle
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::shuffle(order.begin(), order.end(), rng);
    s.type.assign(n, 1);
    for (int r = 0; r < n; ++r) s.type[order[r]] = 1 + (int)((long long)r * ntypes / n);
    s.tag.resize(n);
    for (int i = 0; i < n; ++i) s.tag[i] = i + 1;
    s.mass.resize(n);
    for (int i = 0; i < n; ++i) s.mass[i] = type_mass[s.type[i] - 1];
    // Maxwell-Boltzmann velocities, zero momentum
    s.v.assign(3 * (size_t)n, 0.0);
    double P[3] = {0, 0, 0}, M = 0;
    for (int i = 0; i < n; ++i) {
        std::normal_distribution<double> nd(0.0, std::sqrt(kBoltz * T0 / (s.mass[i] * kMvv2e)));
        for (int q = 0; q < 3; ++q) { s.v[3 * i + q] = nd(rng); P[q] += s.mass[i] * s.v[3 * i + q]; }
        M += s.mass[i];
    }
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < 3; ++q) s.v[3 * i + q] -= P[q] / M;

    const double cutoff_skin = 1.0;
    double cutoff = 0.0;
    mtp_ctx* ctx = nullptr;
    {
        // the MTP cutoff argument of the pair style: the potential's max_dist
        int level = 0, sp = 0, amc = 0, B = 0, T = 0, S = 0;
        long long lds = 0;
        if (mtp_check_file(mtp_path, 64, &level, &sp, &amc, &B, &T, &S, &lds) != MTP_OK) {
            std::fprintf(stderr, "bad potential: %s\n", mtp_global_error());
            return 1;
        }
        FILE* fp = std::fopen(mtp_path, "r");
        char line[4096];
        while (fp && std::fgets(line, sizeof line, fp))
            if (std::strstr(line, "max_dist")) cutoff = std::atof(std::strchr(line, '=') + 1);
        if (fp) std::fclose(fp);
    }
    const int max_nbr = 128;
    if (mtp_open(mtp_path, n, max_nbr, cutoff, 1, &ctx) != MTP_OK) {
        std::fprintf(stderr, "mtp_open: %s\n", mtp_global_error());
        return 1;
    }
    std::vector<double> pos((size_t)n * 3 * max_nbr);
    std::vector<int> nbr((size_t)n * max_nbr), ntype((size_t)n * max_nbr), counts(n), atypes(n), t2d;
    std::vector<const double*> prow(n);
    std::vector<const int*> nrow(n), trow(n);
    for (int i = 0; i < n; ++i) {
        prow[i] = pos.data() + (size_t)i * 3 * max_nbr;
        nrow[i] = nbr.data() + (size_t)i * max_nbr;
        trow[i] = ntype.data() + (size_t)i * max_nbr;
    }
    std::vector<int> ilist(n);
    for (int i = 0; i < n; ++i) ilist[i] = i;
    s.f.assign(3 * (size_t)n, 0.0);
    const double volume = s.L[0] * s.L[1] * s.L[2];

    auto compute = [&](double& pe) -> bool {
        wrap(s);
        make_ghosts(s, cutoff + cutoff_skin);
        std::vector<std::vector<int>> nl;
        neighbor_list(s, cutoff + cutoff_skin, nl);
        std::vector<const double*> xr(s.nall());
        for (int j = 0; j < s.nall(); ++j) xr[j] = &s.x[3 * (size_t)j];
        std::vector<int> numneigh(n);
        std::vector<const int*> first(n);
        for (int i = 0; i < n; ++i) { numneigh[i] = (int)nl[i].size(); first[i] = nl[i].data(); }
        mtp::OrthoMinImage mi{{s.L[0], s.L[1], s.L[2]}, {1, 1, 1}};
        mtp::PackOut po{pos.data(), nbr.data(), ntype.data(), counts.data(), atypes.data()};
        long long trunc = 0;
        const long long tot = mtp::pack_full_list(n, xr.data(), s.type.data(), s.tag.data(), n, ilist.data(),
                                                  numneigh.data(), first.data(), cutoff, n, max_nbr, mi, po, t2d,
                                                  &trunc);
        if (trunc) { std::fprintf(stderr, "neighbor overflow (%lld dropped)\n", trunc); return false; }
        double stress6[6], virial6[6];
        const int rc = mtp_compute(ctx, n, (int)tot, prow.data(), atypes.data(), nrow.data(), counts.data(),
                                   trow.data(), volume, &pe, s.f.data(), stress6, virial6);
        if (rc != MTP_OK) { std::fprintf(stderr, "mtp_compute: %s\n", mtp_last_error(ctx)); return false; }
        return true;
    };
    auto dump = [&](int step) {
        char name[64];
        std::snprintf(name, sizeof name, ".dump%06d", step);
        FILE* fp = std::fopen((out + name).c_str(), "wb");
        std::fwrite(&n, sizeof(int), 1, fp);
        std::fwrite(s.L, sizeof(double), 3, fp);
        std::vector<int> t0(s.type.begin(), s.type.begin() + n);
        std::fwrite(t0.data(), sizeof(int), n, fp);
        std::fwrite(s.x.data(), sizeof(double), 3 * (size_t)n, fp);
        std::fwrite(s.f.data(), sizeof(double), 3 * (size_t)n, fp);
        std::fclose(fp);
    };
    FILE* th = std::fopen((out + ".thermo").c_str(), "w");
    double pe = 0.0;
    if (!compute(pe)) return 1;
    auto kinetic = [&]() {
        double ke = 0;
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < 3; ++q) ke += 0.5 * s.mass[i] * s.v[3 * i + q] * s.v[3 * i + q];
        return ke * kMvv2e;
    };
    double ke = kinetic();
    std::fprintf(th, "%d %.17g %.17g %.17g\n", 0, pe, ke, pe + ke);
    if (dump_every > 0) dump(0);
    for (int step = 1; step <= steps; ++step) {
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < 3; ++q) {
                s.v[3 * i + q] += 0.5 * dt * s.f[3 * i + q] / s.mass[i] * kFtm2v;
                s.x[3 * i + q] += dt * s.v[3 * i + q];
            }
        if (!compute(pe)) return 1;
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < 3; ++q) s.v[3 * i + q] += 0.5 * dt * s.f[3 * i + q] / s.mass[i] * kFtm2v;
        ke = kinetic();
        std::fprintf(th, "%d %.17g %.17g %.17g\n", step, pe, ke, pe + ke);
        if (dump_every > 0 && step % dump_every == 0) dump(step);
        if (step % 100 == 0) { std::printf("step %d etot %.10g\n", step, pe + ke); std::fflush(stdout); }
    }
    std::fclose(th);
    mtp_close(ctx);
    return 0;
}
