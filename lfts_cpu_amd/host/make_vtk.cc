// `lfts-vtk <file> mx my mz [--xfast]`: the reference's tools/make_vtk.sh as a host tool.
// Writes <file>.vtk; with --xfast the values are reordered to VTK's x-fastest order.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "vtk.h"

int main(int argc, char* argv[]) {
    bool xfast = false;
    if (argc == 6 && std::strcmp(argv[5], "--xfast") == 0) xfast = true;
    if (argc != 5 && !(argc == 6 && xfast)) {
        std::cout << "Error -> The following arguments are required..." << std::endl;
        std::cout << "(1) Input file name" << std::endl;
        std::cout << "(2) mx - number of lattice points in x-direction." << std::endl;
        std::cout << "(3) my - number of lattice points in y-direction." << std::endl;
        std::cout << "(4) mz - number of lattice points in z-direction." << std::endl;
        std::cout << "(5) optional --xfast: reorder the values to VTK's x-fastest order" << std::endl;
        return 1;
    }
    const int m[3] = {std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
    const long M = (long)m[0] * m[1] * m[2];
    std::cout << "mx=" << m[0] << "\nmy=" << m[1] << "\nmz=" << m[2] << "\nM=" << M << "\ndx=0.1\ndy=0.1\ndz=0.1"
              << std::endl;
    const std::string in(argv[1]);
    auto tok = lfts_vtk::read_first_tokens(in, M);
    if (xfast && (long)tok.size() < M) {
        std::cout << "Error -> " << in << " holds fewer than M values" << std::endl;
        return 1;
    }
    if (!lfts_vtk::write_vtk_tokens(in + ".vtk", tok, m, xfast)) {
        std::cout << "Error -> cannot write " << in << ".vtk" << std::endl;
        return 1;
    }
    return 0;
}
