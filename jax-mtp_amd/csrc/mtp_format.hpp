// .mtp potential file model and text parser (the MLIP format).
//
// The format and key set follow the reference reader mtp.py:14-70
// (read_mtp, _parse_radial_coeffs) and writer mtp.py:93-139 (write_mtp).
// The level is inferred from alpha_moments_count as in conversion.py:14-29.
#pragma once
#include <string>
#include <vector>

namespace mtp {

struct MtpFile {
    std::string version, potential_name, potential_tag, radial_basis_type;
    double scaling = 1.0;
    int species_count = 0;
    double min_dist = 0.0, max_dist = 0.0;
    int radial_basis_size = 0;   // N (Chebyshev terms)
    int radial_funcs_count = 0;  // R
    std::vector<double> radial_coeffs;  // [Sp][Sp][R][N]
    int alpha_moments_count = 0;
    std::vector<int> alpha_index_basic;  // [B][4] (mu, a, b, c)
    std::vector<int> alpha_index_times;  // [T][4] (i1, i2, mult, i3)
    std::vector<int> alpha_moment_mapping;  // [S]
    std::vector<double> species_coeffs;  // [Sp]
    std::vector<double> moment_coeffs;   // [S]
    bool has_scaling = false, has_radial = false, has_species = false, has_moment = false;

    int B() const { return (int)alpha_index_basic.size() / 4; }
    int T() const { return (int)alpha_index_times.size() / 4; }
    int S() const { return (int)alpha_moment_mapping.size(); }
};

// Level from the moment count (conversion.py:14-29); -1 if unknown.
int level_from_moments_count(int amc);

// Parse .mtp text.  Returns false and sets err on malformed input.
bool parse_mtp_text(const std::string& text, MtpFile& out, std::string& err);
bool read_mtp_file(const std::string& path, MtpFile& out, std::string& err);

// Structural checks used before planning: table bounds, topological order
// of alpha_index_times, unique mapping.  Returns false with a message.
bool validate_mtp(const MtpFile& f, std::string& err);

}  // namespace mtp
