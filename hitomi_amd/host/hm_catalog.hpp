// hm_catalog.hpp -- periodic-box particle input for the drop-in program
// (ParticleGadgetClass, particleGadget.hpp:8-655): Gadget-1 snapshots,
// Rockstar halo lists and ASCII randoms, the RSD shift and the periodic wrap.
// Positions are kept as an (N, 3) FP64 array, ready to upload.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace hmhost {

// Gadget-1 header, 256 bytes (particleGadget.hpp:35-51)
struct GadgetHeader {
    unsigned int npart[6];
    double mass[6];
    double time;
    double redshift;
    int flag_sfr;
    int flag_feedback;
    unsigned int npartTotal[6];
    int flag_cooling;
    int num_files;
    double BoxSize;
    double Omega0;
    double OmegaLambda;
    double HubbleParam;
    char fill[256 - 6 * 4 - 6 * 8 - 2 * 8 - 2 * 4 - 6 * 4 - 2 * 4 - 4 * 8];
};
static_assert(sizeof(GadgetHeader) == 256, "Gadget-1 header must be 256 bytes");

struct Catalog {
    std::vector<double> pos;  // n x 3
    std::vector<double> vel;  // n x 3
    long long n_tot = 0;
    double pos_min[3] = {0, 0, 0}, pos_max[3] = {0, 0, 0};
    double redshift = 0, scaleFactor = 0, boxsize = 0, omegaM0 = 0, omegaLambda = 0, hubble = 0, mass = 0;
    int numFiles = 0;

    int init(long long num) {
        if (num <= 0) {
            printf("The number of particles should be > 0\n");
            return -1;
        }
        n_tot = num;
        pos.assign(3 * (size_t)num, 0.0);
        vel.assign(3 * (size_t)num, 0.0);
        return 0;
    }

    // particleGadget.hpp:128-169
    int read_gadget_header(const std::string& base, GadgetHeader& h) {
        std::string f = base + ".0";
        FILE* fp = fopen(f.c_str(), "rb");
        if (!fp) {
            printf("can not open %s\n", f.c_str());
            return -1;
        }
        int dummy = 0;
        size_t ok = fread(&dummy, sizeof(dummy), 1, fp);
        ok += fread(&h, sizeof(h), 1, fp);
        ok += fread(&dummy, sizeof(dummy), 1, fp);
        fclose(fp);
        if (ok != 3) {
            printf("can not read the header of %s\n", f.c_str());
            return -1;
        }
        long long nsum = 0;
        for (int k = 0; k < 6; k++) nsum += h.npartTotal[k];
        n_tot = nsum;
        boxsize = h.BoxSize;
        omegaM0 = h.Omega0;
        omegaLambda = h.OmegaLambda;
        hubble = h.HubbleParam;
        redshift = h.redshift;
        scaleFactor = h.time;
        numFiles = h.num_files;
        double ms = 0.0;
        for (int k = 0; k < 6; k++) ms += h.mass[k];
        mass = ms;
        return 0;
    }

    // particleGadget.hpp:211-298: float32 positions, velocities x sqrt(a), int64 ids skipped
    int read_gadget(const std::string& base) {
        GadgetHeader h;
        if (read_gadget_header(base, h)) {
            printf("'this->readParameters(snapshotFileName)' fails.\n");
            return -1;
        }
        if (init(n_tot)) {
            printf("'this->initializeParticle(this->n_tot)' fails.\n");
            return -1;
        }
        long long pc = 0;
        for (int i = 0; i < numFiles; i++) {
            std::string f = base + "." + std::to_string(i);
            FILE* fp = fopen(f.c_str(), "rb");
            if (!fp) {
                printf("can not open %s\n", f.c_str());
                return -1;
            }
            int dummy = 0;
            GadgetHeader hh;
            bool ok = fread(&dummy, 4, 1, fp) == 1 && fread(&hh, sizeof(hh), 1, fp) == 1 && fread(&dummy, 4, 1, fp) == 1;
            long long np = 0;
            for (int k = 0; k < 6; k++) np += hh.npart[k];
            if (!ok || pc + np > n_tot) {
                fclose(fp);
                printf("readParticlesFromSnapshotOfGadget(std::string & snapshotFileName) fails.\n");
                return -1;
            }
            std::vector<float> block(3 * (size_t)np);
            ok = fread(&dummy, 4, 1, fp) == 1 && fread(block.data(), sizeof(float), 3 * np, fp) == (size_t)(3 * np) &&
                 fread(&dummy, 4, 1, fp) == 1;
            for (long long n = 0; ok && n < np; n++)
                for (int k = 0; k < 3; k++) pos[3 * (pc + n) + k] = (double)block[3 * n + k];
            ok = ok && fread(&dummy, 4, 1, fp) == 1 &&
                 fread(block.data(), sizeof(float), 3 * np, fp) == (size_t)(3 * np) && fread(&dummy, 4, 1, fp) == 1;
            for (long long n = 0; ok && n < np; n++)
                for (int k = 0; k < 3; k++) vel[3 * (pc + n) + k] = (double)block[3 * n + k] * std::sqrt(scaleFactor);
            fclose(fp);
            if (!ok) {
                printf("readParticlesFromSnapshotOfGadget(std::string & snapshotFileName) fails.\n");
                return -1;
            }
            pc += np;
        }
        if (pc != n_tot) {
            printf("readParticlesFromSnapshotOfGadget(std::string & snapshotFileName) fails.\n");
            return -1;
        }
        return calc_min_max();
    }

    // particleGadget.hpp:171-209, 300-388
    int read_rockstar(const std::string& fname, double lmin, double lmax) {
        std::ifstream fin(fname.c_str());
        if (!fin) {
            printf("Cannot open '%s'\n", fname.c_str());
            printf("'this->readParametersFromRockstar(fname_in)' fails.\n");
            return -1;
        }
        std::string str;
        char d[256];
        const auto c = std::string::npos;
        while (std::getline(fin, str)) {
            const char* s = str.c_str();
            if (str.find("Om") != c) sscanf(s, "%255s %255s %lg", d, d, &omegaM0);
            if (str.find("Ol") != c) sscanf(s, "%255s %255s %255s %255s %255s %lg", d, d, d, d, d, &omegaLambda);
            if (str.find("h") != c)
                sscanf(s, "%255s %255s %255s %255s %255s %255s %255s %255s %lg", d, d, d, d, d, d, d, d, &hubble);
            if (str.find("Box size") != c) sscanf(s, "%255s %255s %lg", d, d, &boxsize);
            if (str.find("Particle mass") != c) sscanf(s, "%255s %255s %lg", d, d, &mass);
            if (str.find("#a =") != c) sscanf(s, "%255s %255s %lg", d, d, &scaleFactor);
        }
        if (scaleFactor <= 0.0) {
            printf("'this->readParametersFromRockstar(fname_in)' fails.\n");
            return -1;
        }
        redshift = 1.0 / scaleFactor - 1.0;
        mass /= 1.0e10;
        std::vector<double> p, v;
        std::ifstream f2(fname.c_str());
        while (std::getline(f2, str)) {
            double x[34];
            int ID = 0, PID = 0;
            double dm;
            double px, py, pz, vx, vy, vz, m;
            int r = sscanf(str.c_str(),
                           "%d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf "
                           "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %d",
                           &ID, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &px, &py, &pz, &vx, &vy, &vz, &dm, &dm, &dm, &dm,
                           &dm, &m, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &dm, &PID);
            (void)x;
            if (r != 34) continue;
            if ((lmin < std::log10(m)) && (std::log10(m) < lmax)) {
                p.insert(p.end(), {px, py, pz});
                v.insert(v.end(), {vx, vy, vz});
            }
        }
        if (init((long long)(p.size() / 3))) return -1;
        pos = p;
        vel = v;
        return calc_min_max();
    }

    // particleGadget.hpp:533-563 (global plane-parallel RSD along z)
    int set_rsd(const double* L) {
        const double a = scaleFactor;
        const double H = 100.0 * std::sqrt(omegaM0 / std::pow(a, 3) + (1.0 - omegaM0 - omegaLambda) / std::pow(a, 2) +
                                           omegaLambda);
        const double aH = scaleFactor * H;
        if (aH <= 0.0) return -1;
        for (long long i = 0; i < n_tot; i++) pos[3 * i + 2] += vel[3 * i + 2] / aH;
        return periodic(L);
    }

    // particleGadget.hpp:565-591: one shift per axis
    int periodic(const double* L) {
        for (long long p = 0; p < n_tot; p++)
            for (int a = 0; a < 3; a++) {
                double& x = pos[3 * p + a];
                if (x >= L[a]) x -= L[a];
                else if (x < 0.0) x += L[a];
            }
        return calc_min_max();
    }

    int calc_min_max() {
        if (n_tot <= 0) return -1;
        for (int a = 0; a < 3; a++) pos_min[a] = pos_max[a] = pos[a];
        for (long long i = 0; i < n_tot; i++)
            for (int a = 0; a < 3; a++) {
                if (pos_min[a] > pos[3 * i + a]) pos_min[a] = pos[3 * i + a];
                if (pos_max[a] < pos[3 * i + a]) pos_max[a] = pos[3 * i + a];
            }
        return 0;
    }

    // particleGadget.hpp:450-489
    // particleGadget.hpp:390-448: "x y z" per line, lines that do not parse skipped
    int read_randoms(const std::string& fname) {
        std::ifstream fin(fname.c_str());
        if (fin.fail()) {
            printf("can not open file '%s'...\n", fname.c_str());
            return -1;
        }
        std::string str;
        long long n = 0;
        double x, y, z;
        while (std::getline(fin, str))
            if (sscanf(str.c_str(), "%lf %lf %lf", &x, &y, &z) == 3) n++;
        if (init(n)) return -1;
        fin.clear();
        fin.seekg(0);
        long long i = 0;
        while (std::getline(fin, str) && i < n) {
            if (sscanf(str.c_str(), "%lf %lf %lf", &x, &y, &z) != 3) continue;
            pos[3 * i] = x;
            pos[3 * i + 1] = y;
            pos[3 * i + 2] = z;
            i++;
        }
        calc_min_max();
        return 0;
    }

    // particleGadget.hpp:487-530
    static int print_parameters_recon(const std::string& outdir, const Catalog& D, const Catalog& R, double alpha) {
        std::string path = outdir + "/simulation_parameters";
        FILE* fp = fopen(path.c_str(), "w");
        if (!fp) {
            printf("output directry '%s' does not exist\n", outdir.c_str());
            return -1;
        }
        fprintf(fp, "\n");
        fprintf(fp, "P_D.n_tot = %lld\n", D.n_tot);
        fprintf(fp, "P_R.n_tot = %lld\n", R.n_tot);
        fprintf(fp, "\n");
        fprintf(fp, "P_D.redshift = %.5f\n", D.redshift);
        fprintf(fp, "P_D.scaleFactor = %.5f\n", D.scaleFactor);
        fprintf(fp, "P_D.boxsize [Mpc/h] = %.5f\n", D.boxsize);
        fprintf(fp, "P_D.omegaM0 = %.5f\n", D.omegaM0);
        fprintf(fp, "P_D.omegaLambda = %.5f\n", D.omegaLambda);
        fprintf(fp, "P_D.hubble = %.5f\n", D.hubble);
        fprintf(fp, "P_D.mass [10^10 Msun/h] = %.10f\n", D.mass);
        fprintf(fp, "\n");
        fprintf(fp, "alpha = %.5f\n", alpha);
        fprintf(fp, "nbar / 1.0e-4 [(Mpc/h)^{-3}] = %.5f\n", double(D.n_tot) / std::pow(D.boxsize, 3) / 1.0e-4);
        fprintf(fp, "\n");
        for (int a = 0; a < 3; a++) fprintf(fp, "P_D.pos_min[%d] [Mpc/h] = %.5f\n", a, D.pos_min[a]);
        for (int a = 0; a < 3; a++) fprintf(fp, "P_R.pos_min[%d] [Mpc/h] = %.5f\n", a, R.pos_min[a]);
        fprintf(fp, "\n");
        for (int a = 0; a < 3; a++) fprintf(fp, "P_D.pos_max[%d] [Mpc/h] = %.5f\n", a, D.pos_max[a]);
        for (int a = 0; a < 3; a++) fprintf(fp, "P_R.pos_max[%d] [Mpc/h] = %.5f\n", a, R.pos_max[a]);
        fprintf(fp, "\n");
        fclose(fp);
        return 0;
    }

    int print_parameters(const std::string& outdir) const {
        std::string path = outdir + "/simulation_parameters";
        FILE* fp = fopen(path.c_str(), "w");
        if (!fp) {
            printf("output directry '%s' does not exist\n", outdir.c_str());
            return -1;
        }
        fprintf(fp, "\n");
        fprintf(fp, "P.n_tot = %lld\n", n_tot);
        fprintf(fp, "\n");
        fprintf(fp, "P.redshift = %.5f\n", redshift);
        fprintf(fp, "P.scaleFactor = %.5f\n", scaleFactor);
        fprintf(fp, "P.boxsize [Mpc/h] = %.5f\n", boxsize);
        fprintf(fp, "P.omegaM0 = %.5f\n", omegaM0);
        fprintf(fp, "P.omegaLambda = %.5f\n", omegaLambda);
        fprintf(fp, "P.hubble = %.5f\n", hubble);
        fprintf(fp, "P.mass [10^10 Msun/h] = %.10f\n", mass);
        fprintf(fp, "\n");
        fprintf(fp, "nbar / 1.0e-4 [(Mpc/h)^{-3}] = %.5f\n", double(n_tot) / std::pow(boxsize, 3) / 1.0e-4);
        fprintf(fp, "\n");
        for (int a = 0; a < 3; a++) fprintf(fp, "P.pos_min[%d] [Mpc/h] = %.5f\n", a, pos_min[a]);
        fprintf(fp, "\n");
        for (int a = 0; a < 3; a++) fprintf(fp, "P.pos_max[%d] [Mpc/h] = %.5f\n", a, pos_max[a]);
        fprintf(fp, "\n");
        fclose(fp);
        return 0;
    }
};

}  // namespace hmhost
