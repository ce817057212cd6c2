// .mtp text parser.  Key handling follows mtp.py:31-70: floats for
// scaling/min_dist/max_dist, integer scalars, {..} lists for the alpha
// tables and coefficients, and a radial_coeffs block of Sp*Sp headers
// ("k0-k1") each followed by R lines of N values (mtp.py:14-28).
#include "mtp_format.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace mtp {

int level_from_moments_count(int amc)
{
    static const int table[][2] = {{1, 2},     {2, 4},     {8, 6},     {18, 8},    {41, 10},
                                   {84, 12},   {174, 14},  {350, 16},  {718, 18},  {1352, 20},
                                   {2621, 22}, {4991, 24}, {9396, 26}, {17366, 28}};
    for (auto& e : table)
        if (e[0] == amc) return e[1];
    return -1;
}

static std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Parse every number inside a brace expression, ignoring braces and commas.
static bool parse_numbers(const std::string& s, std::vector<double>& out)
{
    out.clear();
    const char* p = s.c_str();
    while (*p) {
        if (*p == '{' || *p == '}' || *p == ',' || std::isspace((unsigned char)*p)) { ++p; continue; }
        char* end = nullptr;
        double v = std::strtod(p, &end);
        if (end == p) return false;
        out.push_back(v);
        p = end;
    }
    return true;
}

static bool parse_ints(const std::string& s, std::vector<int>& out)
{
    std::vector<double> tmp;
    if (!parse_numbers(s, tmp)) return false;
    out.resize(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) {
        if (tmp[i] != (double)(long long)tmp[i]) return false;
        out[i] = (int)tmp[i];
    }
    return true;
}

static bool is_digits(const std::string& s)
{
    if (s.empty()) return false;
    return std::all_of(s.begin(), s.end(), [](char c) { return std::isdigit((unsigned char)c); });
}

bool parse_mtp_text(const std::string& text, MtpFile& f, std::string& err)
{
    f = MtpFile();
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    int declared_B = -1, declared_T = -1, declared_S = -1;
    while (std::getline(in, line)) {
        ++lineno;
        std::string t = trim(line);
        if (t.empty() || t == "MTP") continue;
        if (t == "radial_coeffs") {
            int sp = f.species_count, R = f.radial_funcs_count, N = f.radial_basis_size;
            if (sp <= 0 || R <= 0 || N <= 0) {
                err = "radial_coeffs block before species_count/radial_funcs_count/radial_basis_size";
                return false;
            }
            f.radial_coeffs.assign((size_t)sp * sp * R * N, 0.0);
            for (int pair = 0; pair < sp * sp; ++pair) {
                if (!std::getline(in, line)) { err = "truncated radial_coeffs block"; return false; }
                ++lineno;
                for (int m = 0; m < R; ++m) {
                    if (!std::getline(in, line)) { err = "truncated radial_coeffs block"; return false; }
                    ++lineno;
                    std::vector<double> v;
                    if (!parse_numbers(trim(line), v) || (int)v.size() != N) {
                        err = "radial_coeffs line " + std::to_string(lineno) + ": expected " + std::to_string(N) +
                              " values";
                        return false;
                    }
                    std::copy(v.begin(), v.end(), f.radial_coeffs.begin() + ((size_t)pair * R + m) * N);
                }
            }
            f.has_radial = true;
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos) continue;  // unknown non key=value lines are ignored (mtp.py:33-70)
        std::string key = trim(t.substr(0, eq)), val = trim(t.substr(eq + 1));
        auto bad = [&](const char* what) {
            err = "line " + std::to_string(lineno) + ": bad value for " + key + " (" + what + ")";
            return false;
        };
        if (key == "scaling" || key == "min_dist" || key == "max_dist") {
            char* end = nullptr;
            double v = std::strtod(val.c_str(), &end);
            if (end == val.c_str()) return bad("float");
            if (key == "scaling") { f.scaling = v; f.has_scaling = true; }
            else if (key == "min_dist") f.min_dist = v;
            else f.max_dist = v;
        } else if (key == "alpha_index_basic" || key == "alpha_index_times" || key == "alpha_moment_mapping") {
            std::vector<int> v;
            if (!parse_ints(val, v)) return bad("integer list");
            if (key != "alpha_moment_mapping" && v.size() % 4) return bad("rows of 4");
            if (key == "alpha_index_basic") f.alpha_index_basic = v;
            else if (key == "alpha_index_times") f.alpha_index_times = v;
            else f.alpha_moment_mapping = v;
        } else if (key == "species_coeffs" || key == "moment_coeffs") {
            std::vector<double> v;
            if (!parse_numbers(val, v)) return bad("float list");
            if (key == "species_coeffs") { f.species_coeffs = v; f.has_species = true; }
            else { f.moment_coeffs = v; f.has_moment = true; }
        } else if (is_digits(val)) {
            int v = std::atoi(val.c_str());
            if (key == "species_count") f.species_count = v;
            else if (key == "radial_basis_size") f.radial_basis_size = v;
            else if (key == "radial_funcs_count") f.radial_funcs_count = v;
            else if (key == "alpha_moments_count") f.alpha_moments_count = v;
            else if (key == "alpha_index_basic_count") declared_B = v;
            else if (key == "alpha_index_times_count") declared_T = v;
            else if (key == "alpha_scalar_moments") declared_S = v;
        } else {
            if (key == "version") f.version = val;
            else if (key == "potential_name") f.potential_name = val;
            else if (key == "potential_tag") f.potential_tag = val;
            else if (key == "radial_basis_type") f.radial_basis_type = val;
        }
    }
    if (declared_B >= 0 && declared_B != f.B()) { err = "alpha_index_basic_count does not match the table"; return false; }
    if (declared_T >= 0 && declared_T != f.T()) { err = "alpha_index_times_count does not match the table"; return false; }
    if (declared_S >= 0 && declared_S != f.S()) { err = "alpha_scalar_moments does not match alpha_moment_mapping"; return false; }
    return true;
}

bool read_mtp_file(const std::string& path, MtpFile& out, std::string& err)
{
    std::ifstream fin(path);
    if (!fin) { err = "cannot open potential file: " + path; return false; }
    std::stringstream ss;
    ss << fin.rdbuf();
    if (!parse_mtp_text(ss.str(), out, err)) { err = path + ": " + err; return false; }
    return true;
}

bool validate_mtp(const MtpFile& f, std::string& err)
{
    if (f.species_count <= 0) { err = "species_count must be positive"; return false; }
    if (f.radial_funcs_count <= 0 || f.radial_basis_size <= 0) { err = "radial sizes must be positive"; return false; }
    if (!f.has_radial) { err = "missing radial_coeffs"; return false; }
    // the counts may be declared after the radial block: check the table against the final ones
    const size_t radial_len =
        (size_t)f.species_count * f.species_count * f.radial_funcs_count * f.radial_basis_size;
    if (f.radial_coeffs.size() != radial_len) {
        err = "radial_coeffs holds " + std::to_string(f.radial_coeffs.size()) + " values, species_count^2 * " +
              "radial_funcs_count * radial_basis_size = " + std::to_string(radial_len);
        return false;
    }
    if (!f.has_species || (int)f.species_coeffs.size() != f.species_count) { err = "species_coeffs missing or wrong length"; return false; }
    if (!f.has_moment || (int)f.moment_coeffs.size() != f.S()) { err = "moment_coeffs missing or wrong length"; return false; }
    if (!(f.max_dist > f.min_dist)) { err = "max_dist must exceed min_dist"; return false; }
    int amc = f.alpha_moments_count, B = f.B(), T = f.T();
    if (amc <= 0 || B <= 0 || B > amc) { err = "bad alpha_moments_count / alpha_index_basic"; return false; }
    for (int b = 0; b < B; ++b) {
        const int* e = &f.alpha_index_basic[4 * b];
        if (e[0] < 0 || e[0] >= f.radial_funcs_count || e[1] < 0 || e[2] < 0 || e[3] < 0) {
            err = "alpha_index_basic row " + std::to_string(b) + " out of range";
            return false;
        }
    }
    // topological order: every operand complete before use (file order).
    std::vector<int> last(amc, -1);
    for (int e = 0; e < T; ++e) {
        const int* t = &f.alpha_index_times[4 * e];
        for (int q : {0, 1, 3})
            if (t[q] < 0 || t[q] >= amc) { err = "alpha_index_times row " + std::to_string(e) + " out of range"; return false; }
        if (t[3] < B) { err = "alpha_index_times target is a basic moment"; return false; }
        last[t[3]] = e;
    }
    for (int e = 0; e < T; ++e) {
        const int* t = &f.alpha_index_times[4 * e];
        if (last[t[0]] > e || last[t[1]] > e) {
            err = "alpha_index_times is not in dependency order at row " + std::to_string(e);
            return false;
        }
    }
    for (int s : f.alpha_moment_mapping)
        if (s < 0 || s >= amc) { err = "alpha_moment_mapping out of range"; return false; }
    return true;
}

}  // namespace mtp
