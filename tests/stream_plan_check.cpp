// Host-side check of the contraction streams (mtp_plan.cpp build_items).
//
// Builds the plan of an .mtp file for a workgroup size, executes the entry
// streams of every layer on the CPU exactly as the atom kernel walks them
// (LDS as a byte-addressed double array, one accumulator per lane, each
// lane storing its sum at its own item ends, partial sums combined in their
// fixed order; up to three entries per step sharing an integer multiplier,
// item ends and destinations in the multiplier word), and compares the moments and their
// adjoints with a direct evaluation of alpha_index_times in file order:
//   forward  m[i3] += mult * m[i1] * m[i2]
//   reverse  dE/dm from E = sum_s xi_s m[mapping[s]], entries in reverse order.
// Prints the largest relative differences; exit status 0 if both are below
// 1e-12 and every structural check held.
//
// Build: g++ -O2 -std=c++17 -I<csrc> stream_plan_check.cpp mtp_plan.cpp mtp_format.cpp
// Usage: stream_plan_check <file.mtp> <threads> <max_nbr> [seed] [pair|abs]
// (pair: two rows interleaved in 16-byte slots, operand fields as slot indices;
// abs: one row, operand fields as 8-byte slot indices from the LDS base)
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
    if (argc < 4) return fail("usage: stream_plan_check <file.mtp> <threads> <max_nbr> [seed] [pair|abs]");
    const int nt = std::atoi(argv[2]), K = std::atoi(argv[3]);
    const unsigned seed = argc > 4 ? (unsigned)std::atoi(argv[4]) : 1u;
    const bool paired = argc > 5 && std::string(argv[5]) == "pair";
    const bool abs_slots = argc > 5 && std::string(argv[5]) == "abs";
    const int W = paired ? 2 : 1;  // rows sharing the streams (values per LDS slot)
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
    // a layout of the same shape as the kernel's (in slots): zero, m, dm, part
    const int mlen = std::max(p.n_live, 4 * K * 3 + 2 * g.radial_funcs_count * K);
    const int dlen = std::max({amc, K * 3 * (p.nu_max + 1) + kMaxR * K, amc - p.n_live + nt});
    StreamLayout lay;
    if (paired) {
        lay.unit = 16;
        lay.m_off = 1;
        lay.absolute = true;
    }
    if (abs_slots) lay.absolute = true;
    lay.dm_off = lay.m_off + mlen;
    lay.part_off = lay.dm_off + dlen;
    ItemPlan ip;
    if (!build_items(p, nt, lay, 1 << 20, ip, err)) return fail(err.c_str());
    const int nw = nt / 64;
    const size_t nslot = (size_t)lay.part_off + ip.max_partials + 8;
    std::vector<double> lds(nslot * W, 0.0);
    // slot of an operand field / destination
    auto slot = [&](uint32_t field, bool rev_b) -> size_t {
        size_t s;
        if (lay.absolute) {
            s = field;
        } else {
            if (field % 8) { std::fprintf(stderr, "FAIL: unaligned offset %u\n", field); std::exit(1); }
            s = field / 8 + (rev_b ? (size_t)lay.dm_off : 0);
        }
        if (s >= nslot) { std::fprintf(stderr, "FAIL: LDS slot %zu out of range\n", s); std::exit(1); }
        return s;
    };
    auto dest_slot = [&](uint32_t byte_off) -> size_t {
        if (byte_off % lay.unit || byte_off / lay.unit >= nslot) {
            std::fprintf(stderr, "FAIL: destination %u out of range\n", byte_off);
            std::exit(1);
        }
        return byte_off / lay.unit;
    };

    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> xi(amc, 0.0), coef(S);
    for (int s = 0; s < S; ++s) coef[s] = U(rng), xi[g.alpha_moment_mapping[s]] += coef[s];
    std::vector<std::vector<double>> md(W, std::vector<double>(amc, 0.0)), dd(W);
    for (int r = 0; r < W; ++r) {
        for (int x = 0; x < B; ++x) md[r][x] = U(rng);
        for (int x = 0; x < B; ++x) lds[(size_t)(lay.m_off + x) * W + r] = md[r][x];
        // direct forward
        for (int e = 0; e < T; ++e) {
            const int* t = &g.alpha_index_times[4 * e];
            md[r][t[3]] += (double)t[2] * md[r][t[0]] * md[r][t[1]];
        }
        // direct reverse
        dd[r] = xi;
        for (int e = T - 1; e >= 0; --e) {
            const int* t = &g.alpha_index_times[4 * e];
            dd[r][t[0]] += (double)t[2] * md[r][t[1]] * dd[r][t[3]];
            dd[r][t[1]] += (double)t[2] * md[r][t[0]] * dd[r][t[3]];
        }
    }

    auto run = [&](const StreamPlan& sp, int L, bool rev) -> int {
        for (int w = 0; w < nw; ++w) {
            const I4 wv = sp.wave[(size_t)L * nw + w];
            if (wv.y % kStreamUnroll) return fail("stream length is not a multiple of the unroll");
            std::vector<double> acc(64 * W, 0.0);
            for (int s = wv.x; s < wv.x + wv.y; ++s) {
                const U4* q = &sp.tbl[(size_t)s * 64];
                for (int l = 0; l < 64; ++l) {
                    const bool end = stream_dest(q[l].w) != 0;  // per lane
                    const uint32_t e[3] = {q[l].x, q[l].y, q[l].z};
                    const uint32_t mult = stream_mult(q[l].w);
                    if (mult == 0 && (e[0] | e[1] | e[2]) != 0) return fail("entries in a step without a multiplier");
                    // the kernel issues both operand reads of all three entries
                    size_t sa[3], sb[3];
                    for (int j = 0; j < 3; ++j) sa[j] = slot(e[j] & 0xffffu, false), sb[j] = slot(e[j] >> 16, rev);
                    for (int r = 0; r < W; ++r) {
                        double t = lds[sa[0] * W + r] * lds[sb[0] * W + r];
                        t = std::fma(lds[sa[1] * W + r], lds[sb[1] * W + r], t);
                        t = std::fma(lds[sa[2] * W + r], lds[sb[2] * W + r], t);
                        acc[l * W + r] = std::fma((double)mult, t, acc[l * W + r]);
                    }
                    if (end) {
                        const size_t d = dest_slot(stream_dest(q[l].w));
                        if (d == 0) return fail("an item stores into the zero slot");
                        for (int r = 0; r < W; ++r) {
                            lds[d * W + r] = acc[l * W + r];
                            acc[l * W + r] = 0.0;
                        }
                    }
                }
            }
            for (int l = 0; l < 64 * W; ++l)
                if (acc[l] != 0.0) return fail("a sum left unstored at the end of a stream");
        }
        for (int r = 0; r < W; ++r)
            if (lds[r] != 0.0) return fail("the zero slot was written with a non-zero value");
        const size_t dst = rev ? lay.dm_off : lay.m_off;
        for (int r = 0; r < W; ++r) {
            for (int c = sp.cb_beg[L]; c < sp.cb_beg[L + 1]; ++c) {
                const I4 cb = sp.comb[c];
                double a = rev ? xi[cb.x] : 0.0;
                for (int q = cb.y; q < cb.z; ++q) a += lds[((size_t)lay.part_off + q) * W + r];
                lds[(dst + cb.x) * W + r] = a;
            }
            if (rev) {
                for (int t = sp.bo_beg[L]; t < sp.bo_beg[L + 1]; ++t)
                    lds[(dst + sp.bo[t]) * W + r] = xi[sp.bo[t]];
                for (int t = sp.xa_beg[L]; t < sp.xa_beg[L + 1]; ++t)
                    lds[(dst + sp.xa[t]) * W + r] += xi[sp.xa[t]];
            }
        }
        return 0;
    };

    for (int d = 0; d < p.n_depth; ++d)
        if (run(ip.fwd, d, false)) return 1;
    double ef = 0.0, er = 0.0;
    for (int r = 0; r < W; ++r) {
        double sc = 0.0, e = 0.0;
        for (int x = 0; x < amc; ++x) sc = std::max(sc, std::fabs(md[r][x]));
        for (int x = 0; x < amc; ++x)
            e = std::max(e, std::fabs(lds[(size_t)(lay.m_off + x) * W + r] - md[r][x]));
        ef = std::max(ef, e / sc);
    }
    // the reverse pass reads m[0, n_live) only: keep those, as the kernel does
    for (int li = 0; li <= p.n_depth; ++li)
        if (run(ip.rev, li, true)) return 1;
    for (int r = 0; r < W; ++r) {
        double sc = 0.0, e = 0.0;
        for (int x = 0; x < amc; ++x) sc = std::max(sc, std::fabs(dd[r][x]));
        for (int x = 0; x < amc; ++x)
            e = std::max(e, std::fabs(lds[(size_t)(lay.dm_off + x) * W + r] - dd[r][x]));
        er = std::max(er, e / sc);
    }
    int steps = 0;
    for (const I4& w : ip.fwd.wave) steps += w.y;
    for (const I4& w : ip.rev.wave) steps += w.y;
    // LDS passes per operand read (gfx950: 16 lanes per pass from 64 banks
    // of 4 bytes, so slots whose addresses coincide modulo 256 bytes
    // collide: 32 slots of 8 bytes, 16 of 16 bytes); a pass costs the
    // largest number of distinct slots that collide
    double passes = 0.0, ideal = 0.0;
    {
        const int pass = 16, nbank = 256 / lay.unit;
        for (int dir = 0; dir < 2; ++dir) {
            const StreamPlan& sp = dir ? ip.rev : ip.fwd;
            for (size_t st = 0; st < sp.tbl.size() / 64; ++st) {
                const U4* q = &sp.tbl[st * 64];
                for (int half = 0; half < 3; ++half) {
                    for (int op = 0; op < 2; ++op)
                        for (int l0 = 0; l0 < 64; l0 += pass) {
                            std::vector<std::vector<uint32_t>> banks(nbank);
                            for (int l = l0; l < l0 + pass; ++l) {
                                const uint32_t x = half == 0 ? q[l].x : (half == 1 ? q[l].y : q[l].z);
                                const uint32_t fld = op ? x >> 16 : x & 0xffffu;
                                uint32_t addr = lay.absolute ? fld * lay.unit
                                                             : fld + ((op && dir) ? (uint32_t)lay.dm_off * 8u : 0u);
                                auto& v = banks[(addr / lay.unit) % nbank];
                                if (std::find(v.begin(), v.end(), addr) == v.end()) v.push_back(addr);
                            }
                            size_t mx = 1;
                            for (auto& v : banks) mx = std::max(mx, v.size());
                            passes += (double)mx;
                            ideal += 1.0;
                        }
                }
            }
        }
    }
    std::printf("LDS passes per operand read: %.2f (conflict-free 1)\n", passes / ideal);
    if (std::getenv("STREAM_PLAN_VERBOSE")) {
        for (int dir = 0; dir < 2; ++dir) {
            const StreamPlan& sp = dir ? ip.rev : ip.fwd;
            const int nl = (int)sp.wave.size() / nw;
            for (int L = 0; L < nl; ++L) {
                std::printf("%s layer %d steps per wave:", dir ? "rev" : "fwd", L);
                for (int w = 0; w < nw; ++w) std::printf(" %d", sp.wave[(size_t)L * nw + w].y);
                std::printf(" | comb %d\n", sp.cb_beg[L + 1] - sp.cb_beg[L]);
            }
        }
    }
    std::printf("level %d nt %d rows %d: forward %.3e reverse %.3e partials %d steps %d cost %.1f/%.1f\n", p.level, nt,
                W, ef, er, ip.max_partials, steps, ip.fwd.cost, ip.rev.cost);
    return (ef < 1e-12 && er < 1e-12) ? 0 : 1;
}
