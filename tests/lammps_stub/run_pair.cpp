// Drives the pair style's own source (jax-mtp_amd/lammps/pair_jax_mtp.cpp)
// through the LAMMPS stand-ins of lammps_stub.h, the way an input script
// like the reference's test_lammps.in does:
//
//   read_data <file> ; replicate rx ry rz ; pair_style jax/mtp <file.mtp>
//   MAX_ATOMS MAX_NBR CUTOFF [DEBUG] ; pair_coeff * * ; run <steps> (NVE)
//
// Per force call: periodic ghost copies carrying their owners' tags, a full
// neighbor list over locals + ghosts with a skin, forces cleared, then
// PairJaxMTP::compute(eflag, vflag).  As with "neighbor 1.5 bin" and
// "neigh_modify every 1 delay 5 check yes": the ghosts and the list are
// rebuilt when an atom has moved more than half the skin since the last
// build (at most every 5 steps); in between, ghost positions follow their
// owners (forward communication) and neighbor->ago counts the steps.  Step 0 asks for the global and
// per-atom energy and virial (eflag 3, vflag 5); later steps for the
// global energy.  Velocity Verlet in metal units.
//
// Usage: run_pair <file.mtp> <data file> <rx> <ry> <rz> <max_atoms> <max_nbr> <cutoff> <steps> <dt_ps>
//                 <T_K> <seed> <out_prefix> [dump_every] [mass per type ...]
// Writes <out>.thermo (step pe ke etot), <out>.step0 (binary: n, E, virial[6],
// eatom[n], vatom[n][6]) and <out>.dumpNNNNNN (binary: n, box[3], types,
// positions, forces), the format of md_nve's dumps.  Exit status 0 on
// success; a pair-style error prints its message and exits with 3.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "pair_jax_mtp.h"

using namespace LAMMPS_NS;

namespace {

constexpr double kFtm2v = 1.0 / 1.0364269e-4;  // metal: (eV/A) / (g/mol) -> A/ps^2
constexpr double kMvv2e = 1.0364269e-4;        // metal: (g/mol) A^2/ps^2 -> eV
constexpr double kBoltz = 8.617343e-5;         // eV/K

struct Data {
    double lo[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    int ntypes = 0;
    std::vector<double> x;
    std::vector<int> type;
};

// LAMMPS data file, atom_style atomic, orthogonal box ("id type x y z [ix iy iz]")
bool read_data(const char* path, Data& d, std::string& err)
{
    std::ifstream in(path);
    if (!in) { err = std::string("cannot open ") + path; return false; }
    std::string line;
    long natoms = -1;
    double hi[3] = {0, 0, 0};
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        if (line.find("atom types") != std::string::npos) ss >> d.ntypes;
        else if (line.find("atoms") != std::string::npos) ss >> natoms;
        else if (line.find("xlo xhi") != std::string::npos) ss >> d.lo[0] >> hi[0];
        else if (line.find("ylo yhi") != std::string::npos) ss >> d.lo[1] >> hi[1];
        else if (line.find("zlo zhi") != std::string::npos) ss >> d.lo[2] >> hi[2];
        else if (line.find("xy xz yz") != std::string::npos) {
            double t[3];
            ss >> t[0] >> t[1] >> t[2];
            if (t[0] != 0 || t[1] != 0 || t[2] != 0) { err = "triclinic boxes are not supported here"; return false; }
        } else if (line.rfind("Atoms", 0) == 0) break;
    }
    if (natoms <= 0 || d.ntypes <= 0) { err = "data file header incomplete"; return false; }
    for (int q = 0; q < 3; ++q) d.L[q] = hi[q] - d.lo[q];
    std::vector<std::pair<long, std::array<double, 4>>> rows;
    while ((long)rows.size() < natoms && std::getline(in, line)) {
        std::istringstream ss(line);
        long id;
        int t;
        double x, y, z;
        if (!(ss >> id >> t >> x >> y >> z)) continue;
        rows.push_back({id, {(double)t, x, y, z}});
    }
    if ((long)rows.size() != natoms) { err = "Atoms section incomplete"; return false; }
    std::sort(rows.begin(), rows.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& r : rows) {
        d.type.push_back((int)r.second[0]);
        for (int q = 0; q < 3; ++q) d.x.push_back(r.second[1 + q] - d.lo[q]);
    }
    return true;
}

struct Sys {
    int nlocal = 0;
    double L[3] = {0, 0, 0};
    std::vector<double> x;  // [nall][3], locals first
    std::vector<int> type;
    std::vector<int> tag;
    std::vector<int> owner;     // [nghost] local atom of each ghost
    std::vector<double> shift;  // [nghost][3] its periodic image shift
    std::vector<double> v, f, mass;
    std::vector<double*> xrow, frow;
    int nall() const { return (int)type.size(); }
};

void wrap(Sys& s)
{
    for (int i = 0; i < s.nlocal; ++i)
        for (int q = 0; q < 3; ++q) {
            double& c = s.x[3 * i + q];
            c -= s.L[q] * std::floor(c / s.L[q]);
        }
}

void make_ghosts(Sys& s, double rg)
{
    s.x.resize(3 * (size_t)s.nlocal);
    s.type.resize(s.nlocal);
    s.tag.resize(s.nlocal);
    s.owner.clear();
    s.shift.clear();
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
                    s.owner.push_back(i);
                    const double sh[3] = {a * s.L[0], b * s.L[1], c * s.L[2]};
                    s.shift.insert(s.shift.end(), sh, sh + 3);
                }
}

// ghost positions from their owners (forward communication between rebuilds)
void update_ghosts(Sys& s)
{
    for (size_t g = 0; g < s.owner.size(); ++g)
        for (int q = 0; q < 3; ++q)
            s.x[3 * (s.nlocal + g) + q] = s.x[3 * (size_t)s.owner[g] + q] + s.shift[3 * g + q];
}

void neighbor_list(const Sys& s, double rn, std::vector<std::vector<int>>& nl)
{
    const int nall = s.nall();
    int nb[3];
    double bs[3];
    for (int q = 0; q < 3; ++q) {
        nb[q] = std::max(1, (int)std::floor((s.L[q] + 2 * rn) / rn));
        bs[q] = (s.L[q] + 2 * rn) / nb[q];
    }
    auto bin_of = [&](const double* p, int* b) {
        for (int q = 0; q < 3; ++q) b[q] = std::min(nb[q] - 1, std::max(0, (int)((p[q] + rn) / bs[q])));
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
        std::fprintf(stderr, "usage: run_pair <file.mtp> <data> <rx> <ry> <rz> <max_atoms> <max_nbr> <cutoff> <steps> "
                             "<dt_ps> <T_K> <seed> <out_prefix> [dump_every] [mass per type ...]\n");
        return 2;
    }
    Data d;
    std::string err;
    if (!read_data(argv[2], d, err)) { std::fprintf(stderr, "read_data: %s\n", err.c_str()); return 2; }
    const int rep[3] = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
    const int steps = std::atoi(argv[9]);
    const double dt = std::atof(argv[10]), T0 = std::atof(argv[11]);
    const unsigned long long seed = std::strtoull(argv[12], nullptr, 10);
    const std::string out = argv[13];
    const int dump_every = argc > 14 ? std::atoi(argv[14]) : 0;
    std::vector<double> type_mass;
    for (int q = 15; q < argc; ++q) type_mass.push_back(std::atof(argv[q]));
    if ((int)type_mass.size() < d.ntypes) type_mass.resize(d.ntypes, 1.0);

    // replicate (images in x-fastest order), ids 1..n
    Sys s;
    const int n0 = (int)d.type.size();
    for (int c = 0; c < rep[2]; ++c)
        for (int b = 0; b < rep[1]; ++b)
            for (int a = 0; a < rep[0]; ++a)
                for (int i = 0; i < n0; ++i) {
                    s.x.push_back(d.x[3 * i] + a * d.L[0]);
                    s.x.push_back(d.x[3 * i + 1] + b * d.L[1]);
                    s.x.push_back(d.x[3 * i + 2] + c * d.L[2]);
                    s.type.push_back(d.type[i]);
                }
    s.nlocal = (int)s.type.size();
    const int n = s.nlocal;
    for (int q = 0; q < 3; ++q) s.L[q] = rep[q] * d.L[q];
    s.tag.resize(n);
    for (int i = 0; i < n; ++i) s.tag[i] = i + 1;
    s.mass.resize(n);
    for (int i = 0; i < n; ++i) s.mass[i] = type_mass[s.type[i] - 1];
    std::mt19937_64 rng(seed);
    s.v.assign(3 * (size_t)n, 0.0);
    double P[3] = {0, 0, 0}, M = 0;
    for (int i = 0; i < n; ++i) {
        std::normal_distribution<double> nd(0.0, std::sqrt(kBoltz * T0 / (s.mass[i] * kMvv2e)));
        for (int q = 0; q < 3; ++q) { s.v[3 * i + q] = nd(rng); P[q] += s.mass[i] * s.v[3 * i + q]; }
        M += s.mass[i];
    }
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < 3; ++q) s.v[3 * i + q] -= P[q] / M;

    // the LAMMPS objects the pair style sees
    LAMMPS lmp;
    Atom atom;
    Comm comm;
    Domain domain;
    Error error;
    Force force;
    Memory memory;
    Neighbor neighbor;
    Update update;
    NeighList list;
    lmp.atom = &atom; lmp.comm = &comm; lmp.domain = &domain; lmp.error = &error; lmp.force = &force;
    lmp.memory = &memory; lmp.neighbor = &neighbor; lmp.update = &update;
    neighbor.lists[0] = &list;
    atom.ntypes = d.ntypes;
    domain.xprd = domain.h[0] = s.L[0];
    domain.yprd = domain.h[1] = s.L[1];
    domain.zprd = domain.h[2] = s.L[2];
    const double cutoff = std::atof(argv[8]);
    const double skin = 1.5;  // "neighbor 1.5 bin" of the reference's input script

    std::vector<std::vector<int>> nl;
    std::vector<int> ilist(n), numneigh(n);
    std::vector<int*> firstneigh(n);
    for (int i = 0; i < n; ++i) ilist[i] = i;
    std::vector<double> x_built;
    int rebuilds = 0;
    auto setup = [&](bool first) {
        bool rebuild = first;
        if (!first && neighbor.ago + 1 >= 5) {  // delay 5, check yes
            double dmax2 = 0.0;
            for (int i = 0; i < n; ++i) {
                double d2 = 0.0;
                for (int q = 0; q < 3; ++q) { const double t = s.x[3 * i + q] - x_built[3 * i + q]; d2 += t * t; }
                dmax2 = std::max(dmax2, d2);
            }
            rebuild = dmax2 > 0.25 * skin * skin;
        }
        if (rebuild) {
            wrap(s);
            make_ghosts(s, cutoff + skin);
            neighbor_list(s, cutoff + skin, nl);
            x_built.assign(s.x.begin(), s.x.begin() + 3 * (size_t)n);
            neighbor.ago = 0;
            ++rebuilds;
        } else {
            update_ghosts(s);
            ++neighbor.ago;
        }
        const int nall = s.nall();
        s.f.assign(3 * (size_t)nall, 0.0);  // force_clear
        s.xrow.resize(nall);
        s.frow.resize(nall);
        for (int j = 0; j < nall; ++j) { s.xrow[j] = &s.x[3 * (size_t)j]; s.frow[j] = &s.f[3 * (size_t)j]; }
        for (int i = 0; i < n; ++i) { numneigh[i] = (int)nl[i].size(); firstneigh[i] = nl[i].data(); }
        atom.nlocal = n;
        atom.nghost = nall - n;
        atom.x = s.xrow.data();
        atom.f = s.frow.data();
        atom.type = s.type.data();
        atom.tag = s.tag.data();
        list.inum = n;
        list.ilist = ilist.data();
        list.numneigh = numneigh.data();
        list.firstneigh = firstneigh.data();
    };

    int rc = 0;
    try {
        PairJaxMTP pair(&lmp);
        std::vector<std::string> sargs = {argv[1], argv[6], argv[7], argv[8], "0"};
        std::vector<char*> sa;
        for (auto& a : sargs) sa.push_back(&a[0]);
        pair.settings((int)sa.size(), sa.data());
        std::string star1 = "*", star2 = "*";
        char* ca[2] = {&star1[0], &star2[0]};
        pair.coeff(2, ca);
        pair.init_style();
        if (pair.init_one(1, 1) != cutoff) throw std::runtime_error("init_one does not return the cutoff");
        if (neighbor.requests != 1 || !(neighbor.request_flags & NeighConst::REQ_FULL))
            throw std::runtime_error("the pair style did not request a full neighbor list");

        FILE* th = std::fopen((out + ".thermo").c_str(), "w");
        auto kinetic = [&]() {
            double ke = 0;
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < 3; ++q) ke += 0.5 * s.mass[i] * s.v[3 * i + q] * s.v[3 * i + q];
            return ke * kMvv2e;
        };
        auto dump = [&](int step) {
            char name[64];
            std::snprintf(name, sizeof name, ".dump%06d", step);
            FILE* fp = std::fopen((out + name).c_str(), "wb");
            std::fwrite(&n, sizeof(int), 1, fp);
            std::fwrite(s.L, sizeof(double), 3, fp);
            std::fwrite(s.type.data(), sizeof(int), n, fp);
            std::fwrite(s.x.data(), sizeof(double), 3 * (size_t)n, fp);
            std::fwrite(s.f.data(), sizeof(double), 3 * (size_t)n, fp);
            std::fclose(fp);
        };
        setup(true);
        pair.compute(3, 5);  // global + per-atom energy and virial
        {
            FILE* fp = std::fopen((out + ".step0").c_str(), "wb");
            std::fwrite(&n, sizeof(int), 1, fp);
            std::fwrite(&pair.eng_vdwl, sizeof(double), 1, fp);
            std::fwrite(pair.virial, sizeof(double), 6, fp);
            std::fwrite(pair.eatom, sizeof(double), n, fp);
            for (int i = 0; i < n; ++i) std::fwrite(pair.vatom[i], sizeof(double), 6, fp);
            std::fclose(fp);
        }
        double pe = pair.eng_vdwl, ke = kinetic();
        std::fprintf(th, "%d %.17g %.17g %.17g\n", 0, pe, ke, pe + ke);
        if (dump_every > 0) dump(0);
        for (int step = 1; step <= steps; ++step) {
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < 3; ++q) {
                    s.v[3 * i + q] += 0.5 * dt * s.f[3 * i + q] / s.mass[i] * kFtm2v;
                    s.x[3 * i + q] += dt * s.v[3 * i + q];
                }
            setup(false);
            update.ntimestep = step;
            pair.compute(1, 0);
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < 3; ++q) s.v[3 * i + q] += 0.5 * dt * s.f[3 * i + q] / s.mass[i] * kFtm2v;
            pe = pair.eng_vdwl;
            ke = kinetic();
            std::fprintf(th, "%d %.17g %.17g %.17g\n", step, pe, ke, pe + ke);
            if (dump_every > 0 && step % dump_every == 0) dump(step);
            if (step % 100 == 0) { std::printf("step %d etot %.10g\n", step, pe + ke); std::fflush(stdout); }
        }
        std::fclose(th);
        std::printf("neighbor list builds: %d\n", rebuilds);
        std::printf("%s", lmp.log.c_str());
        for (auto& w : error.warnings) std::printf("WARNING: %s\n", w.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        rc = 3;
    }
    return rc;
}
