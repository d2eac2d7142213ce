// VTK legacy export of field files for ParaView volume plots (tools/make_vtk.sh:44-64).
//
// The reference's script writes a STRUCTURED_POINTS header (DIMENSIONS mx my mz, spacing
// 0.1) and then the first token of each of the first M lines of the field file, as text.
// write_vtk_tokens reproduces that output byte for byte from the tokens; write_vtk_values
// writes the same file from values in memory (formatted as the field files are, 6
// significant digits), so a driver can export without re-reading its own output.
//
// The field files list points with z fastest, p = (x*my + y)*mz + z
// (src/field_generator.h:18), while VTK's STRUCTURED_POINTS expects x fastest. The
// script's order (kept by default) therefore shows the box with x and z exchanged; with
// xfast = true the values are reordered so that DIMENSIONS mx my mz is the true geometry.
#pragma once
#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace lfts_vtk {

inline void write_header(std::ostream& o, const int* m, long M) {
    o << "# vtk DataFile Version 2.0\n";
    o << "CT Density\n";
    o << "ASCII\n";
    o << "\n";
    o << "DATASET STRUCTURED_POINTS\n";
    o << "DIMENSIONS " << m[0] << " " << m[1] << " " << m[2] << "\n";
    o << "ORIGIN 0.000000 0.000000 0.000000\n";
    o << "SPACING 0.1 0.1 0.1\n";
    o << "\n";
    o << "POINT_DATA " << M << "\n";
    o << "SCALARS scalars float\n";
    o << "LOOKUP_TABLE default\n";
    o << "\n";
}

// index in the field file of the n-th value VTK reads (x fastest)
inline long xfast_source(long n, const int* m) {
    const long x = n % m[0];
    const long y = (n / m[0]) % m[1];
    const long z = n / ((long)m[0] * m[1]);
    return (x * m[1] + y) * m[2] + z;
}

// tokens: the first whitespace-separated token of each line of the field file, in order
inline bool write_vtk_tokens(const std::string& out, const std::vector<std::string>& tokens, const int* m,
                             bool xfast) {
    const long M = (long)m[0] * m[1] * m[2];
    std::ofstream o(out);
    if (!o) return false;
    write_header(o, m, M);
    const long n = std::min<long>(M, (long)tokens.size());
    for (long i = 0; i < n; i++) o << tokens[xfast ? xfast_source(i, m) : i] << "\n";
    return (bool)o;
}

inline bool write_vtk_values(const std::string& out, const double* v, const int* m, bool xfast) {
    const long M = (long)m[0] * m[1] * m[2];
    std::vector<std::string> tok(M);
    std::ostringstream s;
    for (long i = 0; i < M; i++) {
        s.str("");
        s << v[i];
        tok[i] = s.str();
    }
    return write_vtk_tokens(out, tok, m, xfast);
}

// the first token of each line (what the script's awk -F '\t' '{print $1}' keeps)
inline std::vector<std::string> read_first_tokens(const std::string& file, long maxLines) {
    std::vector<std::string> t;
    std::ifstream in(file);
    std::string line;
    while ((long)t.size() < maxLines && std::getline(in, line)) t.push_back(line.substr(0, line.find('\t')));
    return t;
}

}  // namespace lfts_vtk
