// Host-side check of the contraction streams (mtp_plan.cpp build_items).
//
// Builds the plan of an .mtp file for a workgroup size, executes the entry
// streams of every layer on the CPU exactly as the atom kernel walks them
// (LDS as a byte-addressed double array, one accumulator per lane, the
// group end read from lane 0 and required to agree on all lanes, partial
// sums combined in their fixed order), and compares the moments and their
// adjoints with a direct evaluation of alpha_index_times in file order:
//   forward  m[i3] += mult * m[i1] * m[i2]
//   reverse  dE/dm from E = sum_s xi_s m[mapping[s]], entries in reverse order.
// Prints the largest relative differences; exit status 0 if both are below
// 1e-12 and every structural check held.
//
// Build: g++ -O2 -std=c++17 -I<csrc> stream_plan_check.cpp mtp_plan.cpp mtp_format.cpp
// Usage: stream_plan_check <file.mtp> <threads> <max_nbr> [seed]
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mtp_plan.hpp"

using namespace mtp;

static int fail(const char* msg)
{
    std::fprintf(stderr, "FAIL: %s\n", msg);
    return 1;
}

int main(int argc, char** argv)
{
    if (argc < 4) return fail("usage: stream_plan_check <file.mtp> <threads> <max_nbr> [seed]");
    const int nt = std::atoi(argv[2]), K = std::atoi(argv[3]);
    const unsigned seed = argc > 4 ? (unsigned)std::atoi(argv[4]) : 1u;
    MtpFile f;
    std::string err;
    if (!read_mtp_file(argv[1], f, err)) return fail(err.c_str());
    // untrained files carry no coefficients: the streams do not depend on them
    if (!f.has_radial) {
        f.radial_coeffs.assign((size_t)f.species_count * f.species_count * f.radial_funcs_count * f.radial_basis_size,
                               0.0);
        f.has_radial = true;
    }
    if (!f.has_species) f.species_coeffs.assign(f.species_count, 0.0), f.has_species = true;
    if (!f.has_moment) f.moment_coeffs.assign(f.S(), 0.0), f.has_moment = true;
    HostPlan p;
    if (!build_plan(f, 4, p, err)) return fail(err.c_str());
    const MtpFile& g = p.file;  // renumbered
    const int amc = g.alpha_moments_count, B = g.B(), T = g.T(), S = g.S();
    // a layout of the same shape as the kernel's: zero slot, m, dm, part
    const int mlen = std::max(p.n_live, 4 * K * 3 + 2 * g.radial_funcs_count * K);
    const int dlen = std::max({amc, K * 3 * (p.nu_max + 1) + kMaxR * K, amc - p.n_live + nt});
    const int dm_off = kLdsM + mlen, part_off = dm_off + dlen;
    ItemPlan ip;
    if (!build_items(p, nt, dm_off, part_off, 1 << 20, ip, err)) return fail(err.c_str());
    const int nw = nt / 64;
    std::vector<double> lds((size_t)part_off + ip.max_partials + 8, 0.0);
    auto at = [&](uint32_t byte_off) -> double& {
        if (byte_off % 8 || byte_off / 8 >= lds.size()) {
            std::fprintf(stderr, "FAIL: LDS offset %u out of range\n", byte_off);
            std::exit(1);
        }
        return lds[byte_off / 8];
    };

    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> m(amc, 0.0), xi(amc, 0.0), coef(S);
    for (int x = 0; x < B; ++x) m[x] = U(rng);
    for (int s = 0; s < S; ++s) coef[s] = U(rng), xi[g.alpha_moment_mapping[s]] += coef[s];
    // direct forward
    std::vector<double> md = m;
    for (int e = 0; e < T; ++e) {
        const int* t = &g.alpha_index_times[4 * e];
        md[t[3]] += (double)t[2] * md[t[0]] * md[t[1]];
    }
    // direct reverse
    std::vector<double> dd = xi;
    for (int e = T - 1; e >= 0; --e) {
        const int* t = &g.alpha_index_times[4 * e];
        dd[t[0]] += (double)t[2] * md[t[1]] * dd[t[3]];
        dd[t[1]] += (double)t[2] * md[t[0]] * dd[t[3]];
    }

    auto run = [&](const StreamPlan& sp, int L, bool rev) -> int {
        for (int w = 0; w < nw; ++w) {
            const I4 wv = sp.wave[(size_t)L * nw + w];
            if (wv.y % kStreamUnroll) return fail("stream length is not a multiple of the unroll");
            std::vector<double> acc(64, 0.0);
            for (int s = wv.x; s < wv.x + wv.y; ++s) {
                const U4* q = &sp.tbl[(size_t)s * 64];
                const bool flush = q[0].w == kStreamFlush;
                for (int l = 0; l < 64; ++l) {
                    if ((q[l].w == kStreamFlush) != flush) return fail("group end differs between lanes");
                    const uint32_t bb = rev ? (uint32_t)dm_off * 8u : 0u;
                    float mf;
                    std::memcpy(&mf, &q[l].y, 4);
                    acc[l] = std::fma((double)mf * at(q[l].x & 0xffffu), at(bb + (q[l].x >> 16)), acc[l]);
                    if (!flush) {
                        std::memcpy(&mf, &q[l].w, 4);
                        acc[l] = std::fma((double)mf * at(q[l].z & 0xffffu), at(bb + (q[l].z >> 16)), acc[l]);
                    }
                }
                if (flush)
                    for (int l = 0; l < 64; ++l) {
                        at(q[l].z) = acc[l];
                        acc[l] = 0.0;
                    }
            }
        }
        if (lds[0] != 0.0) return fail("the zero slot was written with a non-zero value");
        double* dst = rev ? &lds[dm_off] : &lds[kLdsM];
        for (int c = sp.cb_beg[L]; c < sp.cb_beg[L + 1]; ++c) {
            const I4 cb = sp.comb[c];
            double a = rev ? xi[cb.x] : 0.0;
            for (int q = cb.y; q < cb.z; ++q) a += lds[(size_t)part_off + q];
            dst[cb.x] = a;
        }
        if (rev) {
            for (int t = sp.bo_beg[L]; t < sp.bo_beg[L + 1]; ++t) lds[(size_t)dm_off + sp.bo[t]] = xi[sp.bo[t]];
            for (int t = sp.xa_beg[L]; t < sp.xa_beg[L + 1]; ++t) lds[(size_t)dm_off + sp.xa[t]] += xi[sp.xa[t]];
        }
        return 0;
    };

    for (int x = 0; x < B; ++x) lds[(size_t)kLdsM + x] = m[x];
    for (int d = 0; d < p.n_depth; ++d)
        if (run(ip.fwd, d, false)) return 1;
    double ef = 0.0, scale_f = 0.0;
    for (int x = 0; x < amc; ++x) scale_f = std::max(scale_f, std::fabs(md[x]));
    for (int x = 0; x < amc; ++x) ef = std::max(ef, std::fabs(lds[(size_t)kLdsM + x] - md[x]));
    ef /= scale_f;
    // the reverse pass reads m[0, n_live) only: keep those, as the kernel does
    for (int li = 0; li <= p.n_depth; ++li)
        if (run(ip.rev, li, true)) return 1;
    double er = 0.0, scale_r = 0.0;
    for (int x = 0; x < amc; ++x) scale_r = std::max(scale_r, std::fabs(dd[x]));
    for (int x = 0; x < amc; ++x) er = std::max(er, std::fabs(lds[(size_t)dm_off + x] - dd[x]));
    er /= scale_r;
    int steps = 0;
    for (const I4& w : ip.fwd.wave) steps += w.y;
    for (const I4& w : ip.rev.wave) steps += w.y;
    std::printf("level %d nt %d: forward %.3e reverse %.3e partials %d steps %d cost %.1f/%.1f\n", p.level, nt, ef, er,
                ip.max_partials, steps, ip.fwd.cost, ip.rev.cost);
    return (ef < 1e-12 && er < 1e-12) ? 0 : 1;
}
