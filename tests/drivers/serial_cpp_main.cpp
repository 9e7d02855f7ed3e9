// A C++ caller of the system-level interface, shaped like a user program of
// serial_chimes_interface (construct with the replication flag,
// init_chimesFF, calculate on a configuration read from an .xyz file whose
// comment line holds the cell vectors a, b, c).  It compiles against
// include/serial_chimes_interface.h and links libchimescalc.so only.
//
//   serial_cpp_main <parameter file> <xyz file> <small 0/1> [energy seed] [force seed]
//
// The optional seeds are the values energy and forces hold before calculate()
// (force seed s gives atom i component k the value s * (1 + i + 0.25 k)),
// since calculate() adds onto them.  Prints the energy, the number of atom
// type entries after the call, the 9-component stress and the forces at full
// precision.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "serial_chimes_interface.h"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <parameter file> <xyz file> <small 0/1> [energy seed] [force seed]\n", argv[0]);
        return 2;
    }
    const bool small = std::atoi(argv[3]) != 0;
    const double eseed = argc > 4 ? std::atof(argv[4]) : 0.0;
    const double fseed = argc > 5 ? std::atof(argv[5]) : 0.0;
    std::ifstream in(argv[2]);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", argv[2]);
        return 2;
    }
    std::string line;
    std::getline(in, line);
    const int n = std::atoi(line.c_str());
    std::getline(in, line);
    std::vector<double> ca(3), cb(3), cc(3);
    {
        std::istringstream s(line);
        s >> ca[0] >> ca[1] >> ca[2] >> cb[0] >> cb[1] >> cb[2] >> cc[0] >> cc[1] >> cc[2];
    }
    std::vector<std::string> types(n);
    std::vector<double> x(n), y(n), z(n);
    for (int i = 0; i < n; i++) {
        std::getline(in, line);
        std::istringstream s(line);
        s >> types[i] >> x[i] >> y[i] >> z[i];
    }
    double energy = eseed;
    std::vector<double> stress(9, 0.0);
    std::vector<std::vector<double> > force(n, std::vector<double>(3, 0.0));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) force[i][k] = fseed * (1.0 + i + 0.25 * k);

    serial_chimes_interface chimes(small);
    chimes.init_chimesFF(argv[1], 0);
    chimes.calculate(x, y, z, ca, cb, cc, types, energy, force, stress);

    std::printf("energy %.17g\n", energy);
    std::printf("ntypes %zu\n", types.size());
    std::printf("stress");
    for (int k = 0; k < 9; k++) std::printf(" %.17g", stress[k]);
    std::printf("\n");
    for (int i = 0; i < n; i++) std::printf("force %.17g %.17g %.17g\n", force[i][0], force[i][1], force[i][2]);
    return 0;
}
