// Plan construction from the .mtp alpha tables (see mtp_plan.hpp).
#include "mtp_plan.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

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
        for (auto& bl : parts[k]) p.part_blk.push_back(bl);
    }
    for (int k = n_parts; k < 9; ++k) p.part_beg[k] = (int)p.part_blk.size();

    // --- packed entries: (i1 | i2 << 16, mult as float bits) ---------------
    p.fwd_pk.resize(p.fwd_pair.size());
    for (size_t e = 0; e < p.fwd_pair.size(); ++e) {
        float mf = (float)p.fwd_mult[e];
        uint32_t bits;
        std::memcpy(&bits, &mf, 4);
        p.fwd_pk[e] = U2{p.fwd_pair[e], bits};
    }
    p.rev_pk.resize(p.rev_pair.size());
    for (size_t e = 0; e < p.rev_pair.size(); ++e) {
        float mf = (float)p.rev_mult[e];
        uint32_t bits;
        std::memcpy(&bits, &mf, 4);
        p.rev_pk[e] = U2{p.rev_pair[e], bits};
    }

    refresh_plan_params(p);
    return true;
}

// Split one layer (moment positions [l0, l1) with entry ranges beg[] into
// the packed entry array pk, moment ids mom[]) into work items of about
// `chunk` entries: runs of whole moments, or one chunk of a long moment that
// writes a partial sum (partials are combined afterwards in a fixed order).
// Moments without entries go to `base_only`; moments of run items that carry
// a coefficient (`mapped`) go to `xi_add`.  The item tables are transposed,
// entry j of item i at [j * n_items + i], so that the threads of a wave (one
// item each) read consecutive addresses.  Entry word y holds the multiplier
// in its low 16 bits; the last entry of a moment (or chunk) also holds the
// destination moment + 1 (1 for a chunk) in its high 16 bits.
static void split_layer(const std::vector<int>& beg, const std::vector<U2>& pk, const std::vector<int>& mom,
                        const std::vector<char>& mapped, int l0, int l1, int chunk, LayerItems& L,
                        std::vector<U2>& table, int& npart)
{
    npart = 0;
    std::vector<std::vector<U2>> lists;
    auto entry = [&](int e, int dest_plus1) {
        U2 q = pk[e];
        float mf;
        std::memcpy(&mf, &q.y, 4);
        q.y = (uint32_t)mf | ((uint32_t)dest_plus1 << 16);
        return q;
    };
    int pos = l0;
    while (pos < l1) {
        const int n = beg[pos + 1] - beg[pos];
        if (n == 0) { L.base_only.push_back(pos); ++pos; continue; }
        if (n > chunk) {
            const int k = (n + chunk - 1) / chunk;
            const int c = (n + k - 1) / k;
            const int p0 = npart;
            for (int q = 0; q < k; ++q) {
                const int e0 = beg[pos] + q * c, e1 = std::min(beg[pos + 1], e0 + c);
                std::vector<U2> lst;
                for (int e = e0; e < e1; ++e) lst.push_back(entry(e, e == e1 - 1 ? 1 : 0));
                L.items.push_back(I4{(int)lst.size(), pos, npart++, 0});
                lists.push_back(std::move(lst));
            }
            L.comb.push_back(I4{pos, p0, npart, 0});
            ++pos;
            continue;
        }
        const int start = pos;
        int cost = 0;
        std::vector<U2> lst;
        while (pos < l1) {
            const int m = beg[pos + 1] - beg[pos];
            if (m == 0 || m > chunk || (pos > start && cost + m > chunk)) break;
            for (int e = beg[pos]; e < beg[pos + 1]; ++e)
                lst.push_back(entry(e, e == beg[pos + 1] - 1 ? mom[pos] + 1 : 0));
            if (mapped[mom[pos]]) L.xi_add.push_back(mom[pos]);
            cost += m;
            ++pos;
        }
        L.items.push_back(I4{(int)lst.size(), start, -1, 0});
        lists.push_back(std::move(lst));
    }
    size_t maxlen = 0;
    for (auto& l : lists) maxlen = std::max(maxlen, l.size());
    L.table_base = (int)table.size();
    const size_t ni = lists.size();
    table.resize(table.size() + maxlen * ni, U2{0, 0});
    for (size_t i = 0; i < ni; ++i)
        for (size_t j = 0; j < lists[i].size(); ++j) table[L.table_base + j * ni + i] = lists[i][j];
}

// Smallest chunk (>= ceil(entries / nt)) whose split yields at most nt items,
// so that every thread of the workgroup gets one item per layer.
static void split_layer_fit(const std::vector<int>& beg, const std::vector<U2>& pk, const std::vector<int>& mom,
                            const std::vector<char>& mapped, int l0, int l1, int nt, LayerItems& L,
                            std::vector<U2>& table, int& npart)
{
    const int ent = beg[l1] - beg[l0];
    int chunk = std::max(4, (ent + nt - 1) / nt);
    const size_t base = table.size();
    for (;;) {
        L = LayerItems();
        table.resize(base);
        split_layer(beg, pk, mom, mapped, l0, l1, chunk, L, table, npart);
        if ((int)L.items.size() <= nt || chunk >= ent) return;
        chunk += std::max(1, chunk / 32);
    }
}

void build_items(const HostPlan& p, int nt, ItemPlan& ip)
{
    ip = ItemPlan();
    ip.nt = nt;
    ip.max_partials = 0;
    const int amc = p.file.alpha_moments_count;
    std::vector<char> mapped(amc, 0), no_map(amc, 0);
    for (int x : p.file.alpha_moment_mapping) mapped[x] = 1;
    for (int d = 1; d <= p.n_depth; ++d) {
        LayerItems L;
        int np = 0;
        split_layer_fit(p.fwd_beg, p.fwd_pk, p.fwd_tgt, no_map, p.fwd_layer[d], p.fwd_layer[d + 1], nt, L,
                        ip.fwd_table, np);
        ip.max_partials = std::max(ip.max_partials, np);
        ip.fwd.push_back(std::move(L));
    }
    for (int li = 0; li <= p.n_depth; ++li) {
        LayerItems L;
        int np = 0;
        split_layer_fit(p.rev_beg, p.rev_pk, p.rev_x, mapped, p.rev_layer[li], p.rev_layer[li + 1], nt, L,
                        ip.rev_table, np);
        ip.max_partials = std::max(ip.max_partials, np);
        ip.rev.push_back(std::move(L));
    }
}

}  // namespace mtp
