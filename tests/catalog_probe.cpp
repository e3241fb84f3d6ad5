// catalog_probe.cpp -- test harness: reads a catalogue with the drop-in's
// host reader (hitomi_amd/host/hm_catalog.hpp), optionally applies the RSD
// shift and prints what the program would measure.  Built and run by
// tests/test_host_catalog.py on the CPU.
//   catalog_probe gadget <base> <rsd 0|1>
//   catalog_probe rockstar <file> <log10_Mmin> <log10_Mmax> <rsd 0|1>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../hitomi_amd/host/hm_catalog.hpp"

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    hmhost::Catalog c;
    const std::string mode = argv[1];
    int rsd = 0;
    if (mode == "gadget") {
        if (c.read_gadget(argv[2])) return 1;
        rsd = std::atoi(argv[3]);
    } else if (mode == "rockstar") {
        if (argc < 6) return 2;
        if (c.read_rockstar(argv[2], std::atof(argv[3]), std::atof(argv[4]))) return 1;
        rsd = std::atoi(argv[5]);
    } else {
        return 2;
    }
    const double L[3] = {c.boxsize, c.boxsize, c.boxsize};
    if (rsd && c.set_rsd(L)) return 1;
    printf("%lld %.17g %.17g %.17g %.17g %.17g\n", c.n_tot, c.scaleFactor, c.omegaM0, c.omegaLambda, c.hubble,
           c.boxsize);
    for (long long i = 0; i < c.n_tot; i++)
        printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", c.pos[3 * i], c.pos[3 * i + 1], c.pos[3 * i + 2], c.vel[3 * i],
               c.vel[3 * i + 1], c.vel[3 * i + 2]);
    return 0;
}
