// Text I/O of the reference's input and output files (src/file_IO.h:13-37): values are
// read with operator>>, written one per line with the stream's default 6 significant
// digits, so output files are byte-compatible with the reference.
#pragma once
#include <fstream>
#include <limits>
#include <string>

namespace file_IO {

inline bool isValidFile(const std::string& fileName) {
    if (fileName.empty()) return false;
    std::ifstream in(fileName);
    return !in.fail();
}

inline void readArray(double* arr, const std::string& fileName, long n, int nIgnore = 0) {
    std::ifstream in(fileName);
    for (int i = 0; i < nIgnore; i++) in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    for (long r = 0; r < n; r++) in >> arr[r];
}

inline void saveArray(const double* arr, const std::string& fileName, long n, bool append = false) {
    std::ofstream out;
    if (append) out.open(fileName, std::ios_base::app);
    else out.open(fileName);
    for (long r = 0; r < n; r++) out << arr[r] << std::endl;
}

}  // namespace file_IO
