// Plan construction from the .mtp alpha tables (see mtp_plan.hpp).
#include "mtp_plan.hpp"

#include <array>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>

namespace mtp {

void refresh_plan_params(HostPlan& p)
{
    const MtpFile& f = p.file;
    p.xi_of.assign(f.alpha_moments_count, 0.0);
    for (int s = 0; s < f.S(); ++s) p.xi_of[f.alpha_moment_mapping[s]] += f.moment_coeffs[s];
}

// Renumber the non-basic moments so that those never read as an operand
// come last.  Their values are needed only for the energy and the basis
// values (phase D), so the balanced-item kernels keep them in LDS that the
// adjoint pass reuses later.  Entry order and the dependency order are
// unchanged; basic moments keep their indices.  Returns the number of
// moments that are basic or read as an operand.
static int renumber_never_operands(MtpFile& f)
{
    const int amc = f.alpha_moments_count, B = f.B(), T = f.T();
    std::vector<char> operand(amc, 0);
    for (int e = 0; e < T; ++e) {
        operand[f.alpha_index_times[4 * e]] = 1;
        operand[f.alpha_index_times[4 * e + 1]] = 1;
    }
    std::vector<int> perm(amc);
    int next = 0;
    for (int x = 0; x < B; ++x) perm[x] = next++;
    for (int x = B; x < amc; ++x)
        if (operand[x]) perm[x] = next++;
    const int n_live = next;
    for (int x = B; x < amc; ++x)
        if (!operand[x]) perm[x] = next++;
    for (int e = 0; e < T; ++e) {
        int* t = &f.alpha_index_times[4 * e];
        t[0] = perm[t[0]];
        t[1] = perm[t[1]];
        t[3] = perm[t[3]];
    }
    for (int& s : f.alpha_moment_mapping) s = perm[s];
    return n_live;
}

bool build_plan(const MtpFile& f_in, int n_parts, HostPlan& p, std::string& err)
{
    if (!validate_mtp(f_in, err)) return false;
    p = HostPlan();
    p.file = f_in;
    p.n_live = renumber_never_operands(p.file);
    const MtpFile& f = p.file;
    p.level = level_from_moments_count(f.alpha_moments_count);
    const int amc = f.alpha_moments_count, B = f.B(), T = f.T(), R = f.radial_funcs_count;
    if (R > kMaxR) { err = "radial_funcs_count above " + std::to_string(kMaxR) + " is not supported"; return false; }
    if (f.radial_basis_size > kMaxN) { err = "radial_basis_size above " + std::to_string(kMaxN) + " is not supported"; return false; }
    if (amc > 65535) { err = "alpha_moments_count above 65535 is not supported"; return false; }

    // --- inverse of the scalar mapping (per-row coefficient vectors) -------
    {
        std::vector<std::vector<int>> by_x(amc);
        for (int s = 0; s < f.S(); ++s) by_x[f.alpha_moment_mapping[s]].push_back(s);
        p.map_beg.push_back(0);
        for (int x = 0; x < amc; ++x) {
            if (by_x[x].empty()) continue;
            p.map_x.push_back(x);
            p.map_s.insert(p.map_s.end(), by_x[x].begin(), by_x[x].end());
            p.map_beg.push_back((int)p.map_s.size());
        }
    }

    // --- basic moments as canonical (mu, nu) blocks -----------------------
    for (auto& row : p.blk_off) for (int& v : row) v = -1;
    for (int& v : p.mu_max_of_nu) v = -1;
    int b = 0;
    while (b < B) {
        const int* e = &f.alpha_index_basic[4 * b];
        int mu = e[0], nu = e[1] + e[2] + e[3];
        if (nu > kMaxNu) { err = "tensor order above " + std::to_string(kMaxNu) + " is not supported"; return false; }
        int n = (nu + 1) * (nu + 2) / 2;
        if (b + n > B) { err = "alpha_index_basic is not made of complete symmetric tensor blocks"; return false; }
        int q = 0;
        for (int a = nu; a >= 0; --a)
            for (int bb = nu - a; bb >= 0; --bb, ++q) {
                const int* r = &f.alpha_index_basic[4 * (b + q)];
                if (r[0] != mu || r[1] != a || r[2] != bb || r[3] != nu - a - bb) {
                    err = "alpha_index_basic is not in canonical (mu, nu) block order";
                    return false;
                }
            }
        if (p.blk_off[mu][nu] != -1) { err = "alpha_index_basic repeats a (mu, nu) block"; return false; }
        p.blk_off[mu][nu] = b;
        p.mu_max_of_nu[nu] = std::max(p.mu_max_of_nu[nu], mu);
        p.nu_max = std::max(p.nu_max, nu);
        b += n;
    }
    // every mu <= mu_max_of_nu[nu] must have the block (kernels loop mu = 0..mu_max)
    for (int nu = 0; nu <= p.nu_max; ++nu)
        for (int mu = 0; mu <= p.mu_max_of_nu[nu]; ++mu)
            if (p.blk_off[mu][nu] < 0) {
                err = "basic (mu, nu) blocks are not contiguous in mu";
                return false;
            }
    for (int nu = 0; nu <= p.nu_max; ++nu) {
        int q = 0;
        for (int a = nu; a >= 0; --a)
            for (int bb = nu - a; bb >= 0; --bb, ++q) p.mono.push_back({a, bb, nu - a - bb, nu | (q << 8)});
    }
    // top byte: the largest mu of the 64 monomials this one shares a
    // wavefront with (monomial t runs on lane t % 64 of some wave for every
    // workgroup size), so the kernels loop over a wave-uniform mu range
    for (size_t g = 0; g < p.mono.size(); g += 64) {
        int gmax = 0;
        for (size_t t = g; t < std::min(p.mono.size(), g + 64); ++t)
            gmax = std::max(gmax, p.mu_max_of_nu[p.mono[t].w & 0xff]);
        for (size_t t = g; t < std::min(p.mono.size(), g + 64); ++t) p.mono[t].w |= gmax << 24;
    }
    static_assert(kMaxR == 8, "mono_dst holds eight mu per monomial");
    p.mono_dst.assign(p.mono.size() * 2, I4{-1, -1, -1, -1});
    for (size_t t = 0; t < p.mono.size(); ++t) {
        const int nu = p.mono[t].w & 0xff, alpha = (p.mono[t].w >> 8) & 0xffff;
        int* d = &p.mono_dst[2 * t].x;
        for (int mu = 0; mu <= p.mu_max_of_nu[nu]; ++mu) d[mu] = p.blk_off[mu][nu] + alpha;
    }

    // --- basic moments by monomial columns ---------------------------------
    // A task covers the monomials (a, b, c0 + i), i < C, of one column of the
    // monomial table, all mu of each, for the slots of one quarter of the
    // slot pairs (four lanes per task).  Its mu count and length come from a
    // small set of shapes (NMU mu per monomial, C monomials; NMU * C <= 8
    // accumulators), the first monomial's mu count rounded up to a shape; the
    // column's later monomials have fewer or equal mu (every reference level)
    // or start a new task.  Tasks are grouped by shape and padded to 16 per
    // wavefront (dummy tasks store nothing), so a wavefront runs one shape.
    {
        std::vector<std::array<I4, 3>> by_shape[kColShapes];
        const auto mono_q = [](int nu, int a, int bb) {  // canonical index of (a, bb, nu - a - bb) in its block
            int q = 0;
            for (int a2 = nu; a2 > a; --a2) q += nu - a2 + 1;
            return q + (nu - a - bb);
        };
        for (int s = 0; s <= p.nu_max; ++s)
            for (int a = s; a >= 0; --a) {
                const int bb = s - a;
                int c0 = 0;
                while (s + c0 <= p.nu_max) {
                    const int nmu0 = p.mu_max_of_nu[s + c0] + 1;
                    if (nmu0 <= 0) { ++c0; continue; }
                    int sh = 0;
                    while (kColNmu[sh] < nmu0) ++sh;
                    std::array<I4, 3> t;
                    t[0] = I4{a, bb, c0, sh};
                    int* dst = &t[1].x;
                    for (int i = 0; i < 8; ++i) dst[i] = -1;
                    int n = 0;
                    while (n < kColLen[sh] && s + c0 + n <= p.nu_max) {
                        const int nu = s + c0 + n, nmu = p.mu_max_of_nu[nu] + 1;
                        if (nmu > kColNmu[sh] || (nmu <= 0 && n > 0)) break;
                        for (int mu = 0; mu < nmu; ++mu)
                            dst[n * kColNmu[sh] + mu] = p.blk_off[mu][nu] + mono_q(nu, a, bb);
                        ++n;
                    }
                    by_shape[sh].push_back(t);
                    c0 += n;
                }
            }
        p.ctask.clear();
        for (int sh = 0; sh < kColShapes; ++sh) {
            auto& v = by_shape[sh];
            if (v.empty()) continue;
            while (v.size() % kColTasksPerWave) v.push_back({I4{0, 0, 0, sh}, I4{-1, -1, -1, -1}, I4{-1, -1, -1, -1}});
            for (auto& t : v) p.ctask.insert(p.ctask.end(), t.begin(), t.end());
        }
        p.n_cwaves = (int)(p.ctask.size() / 3 / kColTasksPerWave);
        // every basic moment has exactly one task destination
        std::vector<int> seen(B, 0);
        for (size_t t = 0; t < p.ctask.size(); t += 3)
            for (const int d : {p.ctask[t + 1].x, p.ctask[t + 1].y, p.ctask[t + 1].z, p.ctask[t + 1].w,
                                p.ctask[t + 2].x, p.ctask[t + 2].y, p.ctask[t + 2].z, p.ctask[t + 2].w})
                if (d >= 0) {
                    if (d >= B || seen[d]++) { err = "monomial-column tasks do not cover the basic moments once"; return false; }
                }
        for (int x = 0; x < B; ++x)
            if (!seen[x]) { err = "monomial-column tasks do not cover the basic moments once"; return false; }
    }

    // --- dependency depths ------------------------------------------------
    p.depth.assign(amc, 0);
    for (int e = 0; e < T; ++e) {
        const int* t = &f.alpha_index_times[4 * e];
        p.depth[t[3]] = std::max(p.depth[t[3]], 1 + std::max(p.depth[t[0]], p.depth[t[1]]));
    }
    p.n_depth = amc ? *std::max_element(p.depth.begin(), p.depth.end()) : 0;
    if (p.n_depth > kMaxDepth) { err = "contraction depth above " + std::to_string(kMaxDepth); return false; }

    // --- forward layers: targets grouped by depth, entries in file order ---
    std::vector<std::vector<int>> entries_of(amc);
    for (int e = 0; e < T; ++e) entries_of[f.alpha_index_times[4 * e + 3]].push_back(e);
    p.fwd_layer.assign(kMaxDepth + 2, 0);
    p.fwd_beg.push_back(0);
    for (int d = 1; d <= p.n_depth; ++d) {
        p.fwd_layer[d] = (int)p.fwd_tgt.size();
        for (int x = 0; x < amc; ++x) {
            if (p.depth[x] != d || entries_of[x].empty()) continue;
            p.fwd_tgt.push_back(x);
            for (int e : entries_of[x]) {
                const int* t = &f.alpha_index_times[4 * e];
                p.fwd_pair.push_back((uint32_t)t[0] | ((uint32_t)t[1] << 16));
                p.fwd_mult.push_back((double)t[2]);
            }
            p.fwd_beg.push_back((int)p.fwd_pair.size());
        }
    }
    for (int d = p.n_depth + 1; d < kMaxDepth + 2; ++d) p.fwd_layer[d] = (int)p.fwd_tgt.size();

    // --- reverse layers: moments grouped by depth, deepest first ----------
    std::vector<std::vector<std::pair<int, int>>> uses(amc);  // (entry, role)
    for (int e = 0; e < T; ++e) {
        const int* t = &f.alpha_index_times[4 * e];
        uses[t[0]].push_back({e, 0});
        uses[t[1]].push_back({e, 1});
    }
    p.rev_layer.assign(kMaxDepth + 2, 0);
    p.rev_beg.push_back(0);
    int li = 0;
    for (int d = p.n_depth; d >= 0; --d, ++li) {
        p.rev_layer[li] = (int)p.rev_x.size();
        for (int x = 0; x < amc; ++x) {
            if (p.depth[x] != d) continue;
            p.rev_x.push_back(x);
            for (auto& u : uses[x]) {
                const int* t = &f.alpha_index_times[4 * u.first];
                int other = u.second == 0 ? t[1] : t[0];
                p.rev_pair.push_back((uint32_t)other | ((uint32_t)t[3] << 16));
                p.rev_mult.push_back((double)t[2]);
            }
            p.rev_beg.push_back((int)p.rev_pair.size());
        }
    }
    for (; li < kMaxDepth + 2; ++li) p.rev_layer[li] = (int)p.rev_x.size();

    // --- per-neighbor gradient: blocks split into n_parts by cost ----------
    if (n_parts < 1) n_parts = 1;
    if (n_parts > 8) n_parts = 8;
    p.n_parts = n_parts;
    std::vector<std::pair<int, I2>> blocks;  // (cost, (mu, nu))
    for (int mu = 0; mu < kMaxR; ++mu)
        for (int nu = 0; nu <= kMaxNu; ++nu)
            if (p.blk_off[mu][nu] >= 0) blocks.push_back({nu == 0 ? 1 : nu * (nu + 1) / 2 + 2, I2{mu, nu}});
    std::stable_sort(blocks.begin(), blocks.end(), [](auto& a, auto& c) { return a.first > c.first; });
    std::vector<std::vector<I2>> parts(n_parts);
    std::vector<int> load(n_parts, 0);
    for (auto& bl : blocks) {
        int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        parts[k].push_back(bl.second);
        load[k] += bl.first;
    }
    p.part_beg.assign(9, 0);
    for (int k = 0; k < n_parts; ++k) {
        p.part_beg[k] = (int)p.part_blk.size();
        for (auto& bl : parts[k]) p.part_blk.push_back(I2{bl.x | (p.blk_off[bl.x][bl.y] << 8), bl.y});
    }
    for (int k = n_parts; k < 9; ++k) p.part_beg[k] = (int)p.part_blk.size();

    refresh_plan_params(p);
    return true;
}

// ---------------------------------------------------------------------------
// Contraction streams.
//
// A layer's rows (a forward target with its entries, or a reverse moment
// with the entries that read it) are cut into items: a row of at most C
// entries is one item that stores its sum in the row's own moment; a longer
// row becomes k = ceil(n / C) near-equal chunks that store partial sums,
// added afterwards in chunk order.  A row's entries are ordered by
// multiplier first, so a chunk holds few distinct multipliers.
//
// A step carries up to three entries of one lane that share a multiplier
// (the kernel forms mult (a0 b0 + a1 b1 + a2 b2)), so an item takes, per
// multiplier of its entries, ceil(count / 3) steps; its last step carries
// the end flag and the item's destination in the multiplier word.  Every
// lane of a wavefront walks its own queue of items back to back (the end
// flag is per lane), so a lane never waits with zero steps for a longer
// item of another lane: items go, most steps first (ties in row order), to
// the least-loaded lane of the first nwu waves (ties: lowest lane), and a
// wave's stream is its longest lane queue, padded with zero steps to whole
// prefetch batches of kStreamUnroll steps.  C and nwu are chosen per layer
// by a model of the time: the mean of the slowest wave's steps and of the
// steps of all waves (the LDS is shared by the workgroups of a CU), plus
// the partial-sum combination.
// ---------------------------------------------------------------------------
namespace {

struct Item { int len, e0, row, dest, steps; };  // entries ord[e0, e0 + len)

constexpr double kCombSteps = 0.5;     // per partial sum added in the combination

struct LayerChoice {
    std::vector<Item> items;                  // sorted, most steps first
    std::vector<std::vector<int>> lane_items;  // [nw * 64]: the items of each lane, in walking order
    std::vector<int> wave_steps;              // [nw]: the longest lane queue of each wave
    int npart = 0;
    double cost = 0;
};

int batch_steps(int steps)
{
    return (steps + kStreamUnroll - 1) / kStreamUnroll * kStreamUnroll;
}

// Runs of equal multipliers of entries ord[e0, e0 + n) (ordered by
// multiplier), in that order.  Returns the steps.
int item_runs(const std::vector<int>& ord, const std::vector<double>& mult, int e0, int n,
              std::vector<std::pair<int, int>>* runs_out)
{
    std::vector<std::pair<int, int>> runs;  // (first position, length)
    for (int j = 0; j < n;) {
        int k = j + 1;
        while (k < n && mult[ord[e0 + k]] == mult[ord[e0 + j]]) ++k;
        runs.push_back({e0 + j, k - j});
        j = k;
    }
    int steps = 0;
    for (const auto& r : runs) steps += (r.second + 2) / 3;
    if (runs_out) *runs_out = std::move(runs);
    return steps;
}

void cut_items(const std::vector<int>& beg, const std::vector<int>& ord, const std::vector<double>& mult, int l0,
               int l1, int C, LayerChoice& lc, const std::vector<int>& dest_of_row, const StreamLayout& lay,
               std::vector<int>& comb_k)
{
    lc = LayerChoice();
    comb_k.clear();
    for (int pos = l0; pos < l1; ++pos) {
        const int n = beg[pos + 1] - beg[pos];
        if (n == 0) continue;
        if (n <= C) {
            lc.items.push_back(Item{n, beg[pos], pos, dest_of_row[pos - l0], item_runs(ord, mult, beg[pos], n, nullptr)});
            continue;
        }
        const int k = (n + C - 1) / C, c = (n + k - 1) / k;
        for (int q = 0; q < k; ++q) {
            const int e0 = beg[pos] + q * c, e1 = std::min(beg[pos + 1], e0 + c);
            lc.items.push_back(Item{e1 - e0, e0, pos, (lay.part_off + lc.npart) * lay.unit,
                                    item_runs(ord, mult, e0, e1 - e0, nullptr)});
            ++lc.npart;
        }
        comb_k.push_back(k);
    }
    std::stable_sort(lc.items.begin(), lc.items.end(), [](const Item& a, const Item& b) { return a.steps > b.steps; });
}

// Items to the lanes of the first nwu waves, most steps first to the
// least-loaded lane; returns the modelled cost.
double assign_lanes(LayerChoice& lc, int nw, int nwu, int nt, const std::vector<int>& comb_k)
{
    using LoadLane = std::pair<int, int>;
    std::priority_queue<LoadLane, std::vector<LoadLane>, std::greater<LoadLane>> pq;
    for (int l = 0; l < nwu * 64; ++l) pq.push({0, l});
    lc.lane_items.assign((size_t)nw * 64, {});
    std::vector<int> load((size_t)nw * 64, 0);
    for (int i = 0; i < (int)lc.items.size(); ++i) {
        const LoadLane t = pq.top();
        pq.pop();
        lc.lane_items[t.second].push_back(i);
        load[t.second] = t.first + lc.items[i].steps;
        pq.push({load[t.second], t.second});
    }
    lc.wave_steps.assign(nw, 0);
    int mx = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        lc.wave_steps[w] = *std::max_element(load.begin() + (size_t)w * 64, load.begin() + (size_t)(w + 1) * 64);
        const int st = batch_steps(lc.wave_steps[w]);
        mx = std::max(mx, st);
        tot += st;
    }
    std::vector<int> ks = comb_k;
    std::sort(ks.begin(), ks.end(), std::greater<int>());
    std::vector<int> tl(nt, 0);
    for (size_t i = 0; i < ks.size(); ++i) tl[i % nt] += ks[i];
    const int comb_max = ks.empty() ? 0 : *std::max_element(tl.begin(), tl.end());
    lc.cost = 0.5 * mx + 0.5 * (double)tot / nw + kCombSteps * comb_max + (ks.empty() ? 0.0 : 2.0);
    return lc.cost;
}

// LDS byte addresses (from the LDS base) of the two operands of an entry word.
void stream_operand_bytes(uint32_t x, const StreamLayout& lay, bool rev, uint32_t& a, uint32_t& b)
{
    const uint32_t fa = x & 0xffffu, fb = x >> 16;
    if (lay.absolute) {
        a = fa * (uint32_t)lay.unit;
        b = fb * (uint32_t)lay.unit;
    } else {
        a = fa;
        b = fb + (rev ? (uint32_t)(lay.dm_off * lay.unit) : 0u);
    }
}

bool stream_layer(const std::vector<int>& beg, const std::vector<uint32_t>& pair, const std::vector<double>& mult,
                  const std::vector<int>& mom, const std::vector<char>& mapped, int l0, int l1, bool rev, int nt,
                  const StreamLayout& lay, int part_budget, StreamPlan& sp, int L, std::string& err)
{
    const int nw = nt / 64;
    // entry words (operand offsets) and integer multipliers
    const uint32_t ms = (uint32_t)lay.m_off;
    auto entry = [&](int e, uint32_t& x, uint32_t& mw) -> bool {
        const uint32_t i1 = pair[e] & 0xffffu, i2 = pair[e] >> 16;
        uint32_t a, b;
        if (lay.absolute) {
            a = ms + i1;
            b = rev ? (uint32_t)lay.dm_off + i2 : ms + i2;
        } else {
            a = (ms + i1) * (uint32_t)lay.unit;
            b = rev ? i2 * (uint32_t)lay.unit : (ms + i2) * (uint32_t)lay.unit;
        }
        if (a > 0xffffu || b > 0xffffu) return false;
        const double mv = mult[e];
        if (!(mv >= 1.0 && mv <= (double)kStreamMultMax && mv == std::floor(mv))) return false;
        x = a | (b << 16);
        mw = (uint32_t)mv;
        return true;
    };
    std::vector<int> dest_of_row(l1 - l0);
    for (int pos = l0; pos < l1; ++pos)
        dest_of_row[pos - l0] = (rev ? lay.dm_off + mom[pos] : lay.m_off + mom[pos]) * lay.unit;
    // entries of each row ordered by multiplier (ties: entry order)
    std::vector<int> ord(beg.empty() ? 0 : beg[l1 > l0 ? l1 : l0]);
    for (int pos = l0; pos < l1; ++pos) {
        std::iota(ord.begin() + beg[pos], ord.begin() + beg[pos + 1], beg[pos]);
        std::stable_sort(ord.begin() + beg[pos], ord.begin() + beg[pos + 1],
                         [&](int x, int y) { return mult[x] < mult[y]; });
    }

    LayerChoice best, lc;
    bool have = false;
    std::vector<int> comb_k;
    int max_len = 0;
    for (int pos = l0; pos < l1; ++pos) max_len = std::max(max_len, beg[pos + 1] - beg[pos]);
    for (int C = 1; C <= max_len; C += C < 64 ? 1 : (C < 256 ? 4 : 16)) {
        if (C < max_len) {
            long long np = 0;
            for (int pos = l0; pos < l1; ++pos) {
                const int n = beg[pos + 1] - beg[pos];
                if (n > C) np += (n + C - 1) / C;
            }
            if (np > part_budget) continue;
        }
        cut_items(beg, ord, mult, l0, l1, C, lc, dest_of_row, lay, comb_k);
        for (int nwu = 1; nwu <= nw; ++nwu) {
            assign_lanes(lc, nw, nwu, nt, comb_k);
            if (!have || lc.cost < best.cost) { best = lc; have = true; }
        }
    }
    if (!have) {  // no entries in this layer
        cut_items(beg, ord, mult, l0, l1, 1 << 30, best, dest_of_row, lay, comb_k);
        assign_lanes(best, nw, nw, nt, comb_k);
    }
    sp.max_partials = std::max(sp.max_partials, best.npart);
    sp.cost += best.cost;

    // comb / base-only / xi-add lists of this layer (row order)
    sp.cb_beg[L] = (int)sp.comb.size();
    sp.bo_beg[L] = (int)sp.bo.size();
    sp.xa_beg[L] = (int)sp.xa.size();
    // partial slots are numbered in row order (cut_items); a row's range is
    // read back from its items' destinations
    {
        std::vector<int> first(l1 - l0, -1), cnt(l1 - l0, 0);
        for (const Item& it : best.items) {
            if (it.dest < lay.part_off * lay.unit) continue;
            const int pslot = it.dest / lay.unit - lay.part_off;
            int& f0 = first[it.row - l0];
            f0 = f0 < 0 ? pslot : std::min(f0, pslot);
            ++cnt[it.row - l0];
        }
        for (int pos = l0; pos < l1; ++pos) {
            const int n = beg[pos + 1] - beg[pos];
            if (n == 0) { sp.bo.push_back(mom[pos]); continue; }
            if (cnt[pos - l0] > 0) {
                sp.comb.push_back(I4{mom[pos], first[pos - l0], first[pos - l0] + cnt[pos - l0], 0});
            } else if (rev && mapped[mom[pos]]) {
                sp.xa.push_back(mom[pos]);
            }
        }
    }

    // streams
    struct StepDesc { int item = -1, run = -1, cnt = 0; };
    const int pass = 16, nbank = 256 / lay.unit;
    for (int w = 0; w < nw; ++w) {
        const int first_step = (int)(sp.tbl.size() / 64);
        const int steps = best.wave_steps[w] > 0 ? batch_steps(best.wave_steps[w]) : 0;
        const size_t t0 = sp.tbl.size();
        sp.tbl.resize(t0 + (size_t)steps * 64, U4{0u, 0u, 0u, 0u});
        auto at = [&](int step, int l) -> U4& { return sp.tbl[t0 + (size_t)step * 64 + l]; };
        // Per lane: the runs of each of its items and, per step, the item,
        // the run it draws from and how many entries.
        std::vector<std::vector<StepDesc>> sd(64, std::vector<StepDesc>(steps));
        std::vector<std::vector<std::vector<std::pair<int, int>>>> runs(64);
        std::vector<std::vector<std::vector<char>>> used(64);
        for (int l = 0; l < 64; ++l) {
            const std::vector<int>& li = best.lane_items[(size_t)w * 64 + l];
            runs[l].resize(li.size());
            used[l].resize(li.size());
            int st = 0;
            for (size_t q = 0; q < li.size(); ++q) {
                const Item& it = best.items[li[q]];
                item_runs(ord, mult, it.e0, it.len, &runs[l][q]);
                used[l][q].assign(it.len, 0);
                for (int r = 0; r < (int)runs[l][q].size(); ++r) {
                    for (int c = runs[l][q][r].second; c > 0;) {
                        StepDesc& d = sd[l][st++];
                        d.item = (int)q;
                        d.run = r;
                        d.cnt = std::min(3, c);
                        c -= d.cnt;
                    }
                }
                const uint32_t doff = (uint32_t)it.dest;
                if (it.dest <= 0 || it.dest % lay.unit != 0 || doff > kStreamDestMask) {
                    err = "a contraction destination does not fit the stream's multiplier word";
                    return false;
                }
                at(st - 1, l).w = doff;
            }
        }
        // Each entry slot of a step is read by two LDS instructions
        // (operand a, operand b) over all lanes.  On gfx950 the LDS serves
        // such a read 16 lanes at a time from 64 banks of 4 bytes, and two
        // of those lanes whose addresses differ but coincide modulo 256
        // bytes (32 slots of 8 bytes, 16 of 16 bytes) serialize (measured:
        // a 64-bit read with lanes at a stride of 16 bytes runs at full
        // rate, at 32 bytes at half rate).  Slot by slot, each lane
        // takes the remaining entry of its step's run whose operands collide
        // least with those already placed in its pass (ties: entry order).
        // This fixes the order in which a lane sums its item: deterministic,
        // as every other order here.
        for (int st = 0; st < steps; ++st) {
            for (int j = 0; j < 3; ++j) {
                for (int p0 = 0; p0 < 64; p0 += pass) {
                    std::vector<std::vector<uint32_t>> ba(nbank), bb(nbank);  // distinct addresses per bank
                    auto cost = [](const std::vector<uint32_t>& v, uint32_t addr) {
                        return std::find(v.begin(), v.end(), addr) != v.end() ? 0 : (int)v.size();
                    };
                    for (int l = p0; l < p0 + pass; ++l) {
                        const StepDesc& d = sd[l][st];
                        if (d.item < 0 || j >= d.cnt) continue;
                        const auto rr = runs[l][d.item][d.run];
                        const Item& it = best.items[best.lane_items[(size_t)w * 64 + l][d.item]];
                        const size_t base = (size_t)(rr.first - it.e0);
                        std::vector<char>& u = used[l][d.item];
                        int pick = -1, best_c = 0;
                        uint32_t pa = 0, pb = 0, px = 0, pm = 0;
                        for (int k = 0; k < rr.second; ++k) {
                            if (u[base + k]) continue;
                            uint32_t x, mw, a, b;
                            if (!entry(ord[rr.first + k], x, mw)) {
                                err = "an operand offset or multiplier does not fit the packed contraction entries";
                                return false;
                            }
                            stream_operand_bytes(x, lay, rev, a, b);
                            const int c = cost(ba[(a / lay.unit) % nbank], a) + cost(bb[(b / lay.unit) % nbank], b);
                            if (pick < 0 || c < best_c) { pick = k; best_c = c; pa = a; pb = b; px = x; pm = mw; }
                            if (c == 0) break;
                        }
                        u[base + pick] = 1;
                        auto add = [](std::vector<uint32_t>& v, uint32_t addr) {
                            if (std::find(v.begin(), v.end(), addr) == v.end()) v.push_back(addr);
                        };
                        add(ba[(pa / lay.unit) % nbank], pa);
                        add(bb[(pb / lay.unit) % nbank], pb);
                        U4& q = at(st, l);
                        (j == 0 ? q.x : (j == 1 ? q.y : q.z)) = px;
                        q.w = (q.w & kStreamDestMask) | (pm << kStreamMultShift);
                    }
                }
            }
        }
        sp.wave.push_back(I4{first_step, steps, 0, 0});
    }
    return true;
}

}  // namespace

bool build_items(const HostPlan& p, int nt, const StreamLayout& lay, int part_budget, ItemPlan& ip,
                 std::string& err)
{
    ip = ItemPlan();
    ip.nt = nt;
    const int amc = p.file.alpha_moments_count;
    std::vector<char> mapped(amc, 0), no_map(amc, 0);
    for (int x : p.file.alpha_moment_mapping) mapped[x] = 1;
    for (int L = 0; L < kMaxDepth + 2; ++L) {
        ip.fwd.cb_beg[L] = ip.fwd.bo_beg[L] = ip.fwd.xa_beg[L] = 0;
        ip.rev.cb_beg[L] = ip.rev.bo_beg[L] = ip.rev.xa_beg[L] = 0;
    }
    for (int d = 1; d <= p.n_depth; ++d)
        if (!stream_layer(p.fwd_beg, p.fwd_pair, p.fwd_mult, p.fwd_tgt, no_map, p.fwd_layer[d], p.fwd_layer[d + 1],
                          false, nt, lay, part_budget, ip.fwd, d - 1, err))
            return false;
    for (int li = 0; li <= p.n_depth; ++li)
        if (!stream_layer(p.rev_beg, p.rev_pair, p.rev_mult, p.rev_x, mapped, p.rev_layer[li], p.rev_layer[li + 1],
                          true, nt, lay, part_budget, ip.rev, li, err))
            return false;
    // end markers of the per-layer list ranges
    for (int L = p.n_depth; L < kMaxDepth + 2; ++L) {
        ip.fwd.cb_beg[L] = (int)ip.fwd.comb.size();
        ip.fwd.bo_beg[L] = (int)ip.fwd.bo.size();
        ip.fwd.xa_beg[L] = (int)ip.fwd.xa.size();
    }
    for (int L = p.n_depth + 1; L < kMaxDepth + 2; ++L) {
        ip.rev.cb_beg[L] = (int)ip.rev.comb.size();
        ip.rev.bo_beg[L] = (int)ip.rev.bo.size();
        ip.rev.xa_beg[L] = (int)ip.rev.xa.size();
    }
    ip.max_partials = std::max(ip.fwd.max_partials, ip.rev.max_partials);
    return true;
}

}  // namespace mtp
