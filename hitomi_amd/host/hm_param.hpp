// hm_param.hpp -- the reference's .ini parameter interface, restated for the
// drop-in programs (ParameterClass, parameter.hpp:8-623).
//
// Same file format and semantics: a line counts only when it has at least
// three whitespace-separated tokens and the second one is "="; keys are found
// by substring search in the whole line (so "NR", "RG" and "ELL" also match
// any line that contains those letters), every key is validated even when the
// measurement does not use it, and data_file / random_file are prefixed with
// data_dir + "/".  Two keys are added, chosen so that no existing key is a
// substring of them and they are no substring of an existing key:
//   precision   = FP32 | FP64   (grids and FFTs; bin sums are FP64 in both)
//   bk_all_rows = True | False  (write bk*_%02d.dat for every ith_kbin and
//                                zeta*_%02d.dat for every ith_rbin)
#pragma once

#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

namespace hmhost {

struct Params {
    double boxsize[3] = {0.0, 0.0, 0.0};
    double volume = 0.0;
    int n_mesh[3] = {0, 0, 0};
    long long n_mesh_tot = 0;
    int ell1 = -1, ell2 = -1, ELL = -1;
    std::string data_file = "None", random_file = "None", output_dir = "None", data_dir = "None",
                data_mpi_file = "None", measure = "None";
    double kmin = 0.0, kmax = 0.0;
    int n_kbin = 1;
    double rmin = 0.0, rmax = 0.0;
    int n_rbin = 1;
    std::string assign = "None";
    int ith_kbin = -1, ith_rbin = -1;
    int NR = -1;
    double b1_fid = 0.0, RG = -1.0;
    std::string flag_recon = "None";
    int n_mesh_recon[3] = {0, 0, 0};
    std::string flag_RSD = "None", sim_data = "None";
    double log10_Mmin = 0.0, log10_Mmax = 0.0;
    int realization = -1;
    // additions
    std::string precision = "FP32";
    std::string bk_all_rows = "False";

    // parameter.hpp:111-254
    int read(const char* path) {
        std::ifstream fin(path);
        if (!fin) {
            printf("cannot open '%s'\n", path);
            return -1;
        }
        std::string str;
        double bx = 0.0, by = 0.0, bz = 0.0;
        int nx = 0, ny = 0, nz = 0, rx = 0, ry = 0, rz = 0;
        char d[256], chk[256];
        char data_file_t[1024] = "None", data_mpi_t[1024] = "None", random_t[1024] = "None", data_dir_t[1024] = "None",
             out_t[1024] = "None", measure_t[1024] = "None", assign_t[1024] = "None", recon_t[1024] = "None",
             rsd_t[1024] = "None", sim_t[1024] = "None", prec_t[1024] = "FP32", rows_t[1024] = "False";
        const auto c = std::string::npos;
        while (std::getline(fin, str)) {
            const char* s = str.data();
            if (sscanf(s, "%255s %255s %255s", d, d, d) != 3) continue;
            sscanf(s, "%255s %255s %255s", d, chk, d);
            if (std::string(chk) != "=") continue;
            if (str.find("realization") != c) sscanf(s, "%255s %255s %d", d, d, &realization);
            if (str.find("boxsize_x") != c) sscanf(s, "%255s %255s %lg", d, d, &bx);
            if (str.find("boxsize_y") != c) sscanf(s, "%255s %255s %lg", d, d, &by);
            if (str.find("boxsize_z") != c) sscanf(s, "%255s %255s %lg", d, d, &bz);
            if (str.find("n_mesh_x") != c) sscanf(s, "%255s %255s %d", d, d, &nx);
            if (str.find("n_mesh_y") != c) sscanf(s, "%255s %255s %d", d, d, &ny);
            if (str.find("n_mesh_z") != c) sscanf(s, "%255s %255s %d", d, d, &nz);
            if (str.find("ell1") != c) sscanf(s, "%255s %255s %d", d, d, &ell1);
            if (str.find("ell2") != c) sscanf(s, "%255s %255s %d", d, d, &ell2);
            if (str.find("ELL") != c) sscanf(s, "%255s %255s %d", d, d, &ELL);
            if (str.find("kmin") != c) sscanf(s, "%255s %255s %lg", d, d, &kmin);
            if (str.find("kmax") != c) sscanf(s, "%255s %255s %lg", d, d, &kmax);
            if (str.find("n_kbin") != c) sscanf(s, "%255s %255s %d", d, d, &n_kbin);
            if (str.find("rmin") != c) sscanf(s, "%255s %255s %lg", d, d, &rmin);
            if (str.find("rmax") != c) sscanf(s, "%255s %255s %lg", d, d, &rmax);
            if (str.find("n_rbin") != c) sscanf(s, "%255s %255s %d", d, d, &n_rbin);
            if (str.find("data_file") != c) sscanf(s, "%255s %255s %1023s", d, d, data_file_t);
            if (str.find("random_file") != c) sscanf(s, "%255s %255s %1023s", d, d, random_t);
            if (str.find("data_dir") != c) sscanf(s, "%255s %255s %1023s", d, d, data_dir_t);
            if (str.find("output_dir") != c) sscanf(s, "%255s %255s %1023s", d, d, out_t);
            if (str.find("data_mpi_file") != c) sscanf(s, "%255s %255s %1023s", d, d, data_mpi_t);
            if (str.find("measure") != c) sscanf(s, "%255s %255s %1023s", d, d, measure_t);
            if (str.find("assignment") != c) sscanf(s, "%255s %255s %1023s", d, d, assign_t);
            if (str.find("ith_kbin") != c) sscanf(s, "%255s %255s %d", d, d, &ith_kbin);
            if (str.find("ith_rbin") != c) sscanf(s, "%255s %255s %d", d, d, &ith_rbin);
            if (str.find("NR") != c) sscanf(s, "%255s %255s %d", d, d, &NR);
            if (str.find("b1_fid") != c) sscanf(s, "%255s %255s %lg", d, d, &b1_fid);
            if (str.find("RG") != c) sscanf(s, "%255s %255s %lg", d, d, &RG);
            if (str.find("flag_recon") != c) sscanf(s, "%255s %255s %1023s", d, d, recon_t);
            if (str.find("n_mesh_recon_x") != c) sscanf(s, "%255s %255s %d", d, d, &rx);
            if (str.find("n_mesh_recon_y") != c) sscanf(s, "%255s %255s %d", d, d, &ry);
            if (str.find("n_mesh_recon_z") != c) sscanf(s, "%255s %255s %d", d, d, &rz);
            if (str.find("flag_RSD") != c) sscanf(s, "%255s %255s %1023s", d, d, rsd_t);
            if (str.find("sim_data") != c) sscanf(s, "%255s %255s %1023s", d, d, sim_t);
            if (str.find("log10_Mmin") != c) sscanf(s, "%255s %255s %lg", d, d, &log10_Mmin);
            if (str.find("log10_Mmax") != c) sscanf(s, "%255s %255s %lg", d, d, &log10_Mmax);
            if (str.find("precision") != c) sscanf(s, "%255s %255s %1023s", d, d, prec_t);
            if (str.find("bk_all_rows") != c) sscanf(s, "%255s %255s %1023s", d, d, rows_t);
        }
        boxsize[0] = bx;
        boxsize[1] = by;
        boxsize[2] = bz;
        n_mesh[0] = nx;
        n_mesh[1] = ny;
        n_mesh[2] = nz;
        n_mesh_recon[0] = rx;
        n_mesh_recon[1] = ry;
        n_mesh_recon[2] = rz;
        data_file = data_file_t;
        data_mpi_file = data_mpi_t;
        random_file = random_t;
        data_dir = data_dir_t;
        output_dir = out_t;
        measure = measure_t;
        assign = assign_t;
        volume = bx * by * bz;
        n_mesh_tot = (long long)nx * ny * nz;
        flag_recon = recon_t;
        flag_RSD = rsd_t;
        sim_data = sim_t;
        precision = prec_t;
        bk_all_rows = rows_t;
        if (check()) return -1;
        data_file = data_dir + "/" + data_file;
        random_file = data_dir + "/" + random_file;
        return 0;
    }

    // parameter.hpp:389-621 (same order, same messages)
    int check() const {
#define HM_REJECT(cond, what, how)              \
    if (cond) {                                 \
        printf("The '%s' is not loaded correctly.\n", what); \
        printf("%s\n", how);                    \
        return -1;                              \
    }
        HM_REJECT(data_file == "None", "data_file", "It should NOT be 'None'.");
        HM_REJECT(random_file == "None", "random_file", "It should NOT be 'None'.");
        HM_REJECT(output_dir == "None", "output_dir", "It should NOT be 'None'.");
        HM_REJECT(data_dir == "None", "data_dir", "It should NOT be 'None'.");
        HM_REJECT(!(measure == "pk" || measure == "bk" || measure == "2PCF" || measure == "3PCF" ||
                    measure == "window2PCF" || measure == "window3PCF"),
                  "measure", "It should be 'pk', 'bk', '2PCF', '3PCF', 'window2PCF', or 'window3PCF'.");
        HM_REJECT(!(boxsize[0] > 0.0 && boxsize[1] > 0.0 && boxsize[2] > 0.0), "boxsize",
                  "It should be boxsize_x > 0, boxsize_y > 0, and boxsize_z > 0.");
        HM_REJECT(!(n_mesh[0] > 0 && n_mesh[1] > 0 && n_mesh[2] > 0), "n_mesh",
                  "It should be n_mesh_x > 0, n_mesh_y > 0, and n_mesh_z > 0.");
        HM_REJECT(!(ell1 >= 0 && ell2 >= 0 && ELL >= 0), "ell1', 'ell2', or 'ELL",
                  "It should be ell1 >= 0, ell2 >= 0, and ELL >= 0.");
        HM_REJECT(!(kmin > 0.0), "kmin", "It should be kmin > 0.0.");
        HM_REJECT(!(kmax > kmin), "kmax", "It should be kmax > kmin.");
        HM_REJECT(!(n_kbin >= 2), "n_kbin", "It should be n_kbin >= 2.");
        HM_REJECT(!(rmin > 0.0), "rmin", "It should be rmin > 0.0.");
        HM_REJECT(!(rmax > rmin), "rmax", "It should be rmax > rmin.");
        HM_REJECT(!(n_rbin >= 2), "n_rbin", "It should be n_rbin >= 2.");
        HM_REJECT(!(assign == "NGP" || assign == "CIC" || assign == "TSC"), "assign", "It should be 'NGP', 'CIC' or 'TSC'");
        HM_REJECT(!(ith_kbin >= 0 && ith_kbin < n_kbin), "ith_kbin", "It should be ith_kbin >= 0 and ith_kbin < n_kbin.");
        HM_REJECT(!(ith_rbin >= 0 && ith_rbin < n_rbin), "ith_rbin", "It should be ith_rbin >= 0 and ith_rbin < n_rbin.");
        HM_REJECT(!(NR >= 0), "NR", "It should be NR >= 0.");
        HM_REJECT(data_mpi_file == "None", "data_mpi_file", "It should NOT be 'None'.");
        HM_REJECT(!(b1_fid > 0.0), "b1_fid", "It should be b1_fid > 0.0.");
        HM_REJECT(!(RG >= 0.0), "RG", "It should be RG >= 0.0.");
        HM_REJECT(!(flag_recon == "True" || flag_recon == "False"), "flag_recon",
                  "It should be flag_recon = True or False.");
        HM_REJECT(!(n_mesh_recon[0] > 0 && n_mesh_recon[1] > 0 && n_mesh_recon[2] > 0), "n_mesh_recon",
                  "It should be n_mesh_recon_x > 0, n_mesh_recon_y > 0, and n_mesh_recon_z > 0.");
        HM_REJECT(!(realization >= 0), "realization", "It should be realization >= 0.");
        HM_REJECT(!(flag_RSD == "True" || flag_RSD == "False"), "flag_RSD", "It should be flag_RSD = True or False.");
        HM_REJECT(!(sim_data == "Gadget" || sim_data == "Rockstar"), "sim_data",
                  "It should be sim_data = Gadget or Rockstar.");
        HM_REJECT(!(log10_Mmin > 0.0), "log10_Mmin", "It should be log10_Mmin > 0.0.");
        HM_REJECT(!(log10_Mmin < log10_Mmax), "log10_Mmax", "It should be log10_Mmax > log10_Mmin.");
        HM_REJECT(!(precision == "FP32" || precision == "FP64"), "precision", "It should be precision = FP32 or FP64.");
        HM_REJECT(!(bk_all_rows == "True" || bk_all_rows == "False"), "bk_all_rows",
                  "It should be bk_all_rows = True or False.");
#undef HM_REJECT
        return 0;
    }

    // parameter.hpp:256-317
    int make_output_directory() {
        struct stat st;
        if (stat(output_dir.c_str(), &st) != 0) {
            if (mkdir(output_dir.c_str(), 0777) == 0) {
                printf("Directry \"%s\" is successfully made.\n ", output_dir.c_str());
            } else {
                printf("Directry \"%s\" fails to be made\n", output_dir.c_str());
                return -1;
            }
        }
        if (realization > 0) {
            char buf[2048];
            snprintf(buf, sizeof(buf), "%s/%04d", output_dir.c_str(), realization);
            if (stat(buf, &st) != 0) {
                if (mkdir(buf, 0777) == 0) {
                    printf("Directry \"%s\" is successfully made.\n ", buf);
                } else {
                    printf("Directry \"%s\" fails to be made\n", buf);
                    return -1;
                }
            }
            output_dir = buf;
        }
        return 0;
    }

    // parameter.hpp:319-387
    int print_parameters() const {
        std::string path = output_dir + "/input_parameters";
        FILE* fp = fopen(path.c_str(), "w");
        if (!fp) {
            printf("output directry '%s' does not exist\n", output_dir.c_str());
            return -1;
        }
        fprintf(fp, "\n");
        fprintf(fp, "data_dir = %s\n", data_dir.c_str());
        fprintf(fp, "data_file = %s\n", data_file.c_str());
        fprintf(fp, "random_file = %s\n", random_file.c_str());
        fprintf(fp, "output_dir = %s\n", output_dir.c_str());
        fprintf(fp, "\n");
        fprintf(fp, "measure = %s\n", measure.c_str());
        fprintf(fp, "\n");
        fprintf(fp, "realization = %d\n", realization);
        fprintf(fp, "\n");
        fprintf(fp, "ell1 = %d\n", ell1);
        fprintf(fp, "ell2 = %d\n", ell2);
        fprintf(fp, "ELL = %d\n", ELL);
        fprintf(fp, "\n");
        fprintf(fp, "boxsize_x = %.5f\n", boxsize[0]);
        fprintf(fp, "boxsize_y = %.5f\n", boxsize[1]);
        fprintf(fp, "boxsize_z = %.5f\n", boxsize[2]);
        fprintf(fp, "\n");
        fprintf(fp, "n_mesh_x = %d\n", n_mesh[0]);
        fprintf(fp, "n_mesh_y = %d\n", n_mesh[1]);
        fprintf(fp, "n_mesh_z = %d\n", n_mesh[2]);
        fprintf(fp, "\n");
        fprintf(fp, "kmin = %.5f\n", kmin);
        fprintf(fp, "kmax = %.5f\n", kmax);
        fprintf(fp, "n_kbin = %d\n", n_kbin);
        fprintf(fp, "\n");
        fprintf(fp, "rmin = %.5f\n", rmin);
        fprintf(fp, "rmax = %.5f\n", rmax);
        fprintf(fp, "n_rbin = %d\n", n_rbin);
        fprintf(fp, "\n");
        fprintf(fp, "flag_recon = %s\n", flag_recon.c_str());
        fprintf(fp, "b1_fid = %.5f\n", b1_fid);
        fprintf(fp, "RG = %.5f\n", RG);
        fprintf(fp, "n_mesh_recon_x = %d\n", n_mesh_recon[0]);
        fprintf(fp, "n_mesh_recon_y = %d\n", n_mesh_recon[1]);
        fprintf(fp, "n_mesh_recon_z = %d\n", n_mesh_recon[2]);
        fprintf(fp, "\n");
        fprintf(fp, "ith_kbin = %d\n", ith_kbin);
        fprintf(fp, "ith_rbin = %d\n", ith_rbin);
        fprintf(fp, "\n");
        fprintf(fp, "sim_data = %s\n", sim_data.c_str());
        fprintf(fp, "flag_RSD = %s\n", flag_RSD.c_str());
        fprintf(fp, "log10_Mmin = %.5f\n", log10_Mmin);
        fprintf(fp, "log10_Mmax = %.5f\n", log10_Mmax);
        fprintf(fp, "\n");
        fprintf(fp, "NR = %d\n", NR);
        fprintf(fp, "data_mpi_file = %s\n", data_mpi_file.c_str());
        fprintf(fp, "\n");
        fprintf(fp, "assignment = %s\n", assign.c_str());
        fprintf(fp, "\n");
        fclose(fp);
        return 0;
    }
};

// tools.hpp:118-132
inline void set_kbin(const Params& p, double* kb) {
    double dk = (p.kmax - p.kmin) / double(p.n_kbin - 1);
    for (int i = 0; i < p.n_kbin; i++) kb[i] = p.kmin + dk * double(i);
}
inline void set_rbin(const Params& p, double* rb) {
    double dr = (p.rmax - p.rmin) / double(p.n_rbin - 1);
    for (int i = 0; i < p.n_rbin; i++) rb[i] = p.rmin + dr * double(i);
}

}  // namespace hmhost
