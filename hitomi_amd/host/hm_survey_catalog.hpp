// hm_survey_catalog.hpp -- survey particle input for the drop-in program
// (ParticleBOSSClass, particle.hpp:8-423): ASCII "x y z w" catalogues read in
// two passes, positions and weights kept as flat FP64 arrays ready to upload.
#pragma once

#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace hmhost {

struct SurveyCatalog {
    std::vector<double> pos;  // n x 3, the file's Cartesian coordinates
    std::vector<double> w;    // n
    long long n_tot = 0;

    // particle.hpp:93-154: lines that do not parse as four numbers are skipped
    int read(const std::string& fname) {
        std::ifstream fin(fname.c_str());
        if (fin.fail()) {
            printf("Cannot open the file '%s'...\n", fname.c_str());
            return -1;
        }
        std::string str;
        long long n = 0;
        double x, y, z, ww;
        while (std::getline(fin, str))
            if (sscanf(str.c_str(), "%lf %lf %lf %lf", &x, &y, &z, &ww) == 4) n++;
        if (n <= 0) {
            printf("The number of particles should be > 0\n");
            printf("'initializeParticle' fails.\n");
            return -1;
        }
        fin.clear();
        fin.seekg(0);
        pos.resize(3 * (size_t)n);
        w.resize((size_t)n);
        long long i = 0;
        while (std::getline(fin, str) && i < n) {
            if (sscanf(str.c_str(), "%lf %lf %lf %lf", &x, &y, &z, &ww) != 4) continue;
            pos[3 * i] = x;
            pos[3 * i + 1] = y;
            pos[3 * i + 2] = z;
            w[i] = ww;
            i++;
        }
        n_tot = n;
        return 0;
    }

    // particle.hpp:160-200 on the positions minus the FFT-box shift
    void min_max_shifted(const double* shift, double* mn, double* mx) const {
        for (int a = 0; a < 3; a++) {
            mn[a] = pos[a] - shift[a];
            mx[a] = mn[a];
        }
        for (long long i = 0; i < n_tot; i++)
            for (int a = 0; a < 3; a++) {
                const double v = pos[3 * i + a] - shift[a];
                if (mn[a] > v) mn[a] = v;
                if (mx[a] < v) mx[a] = v;
            }
    }
};

}  // namespace hmhost
