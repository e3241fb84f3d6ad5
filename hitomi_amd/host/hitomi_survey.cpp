// hitomi_survey -- drop-in for the survey program of hitomi_measurement
// (main.cpp:13-292): `hitomi_survey param.ini`.
//
// Reads the same .ini keys and the same ASCII "x y z w" data and random
// catalogues, writes input_parameters, survey_parameters and pk%d.dat in the
// reference's formats, and exits with status 1 where the reference calls
// EXIT().  The measurement runs on the GPU through libhitomi_amd.so.
//
// Supported: measure = pk, 2PCF and window2PCF (ell1 = ELL, ell2 = 0), and
// bk, 3PCF and window3PCF for any (ell1, ell2, ELL) that wigner_3j allows,
// each with or without reconstruction (flag_recon).
#include "../../include/hitomi_amd.h"
#include "hm_param.hpp"
#include "hm_survey_catalog.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace hmhost;

static double g_t0 = 0.0;
static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static void stage(const char* what) { printf("%s: %.3f sec.\n", what, now() - g_t0); }

[[noreturn]] static void EXIT() { std::exit(1); }

static int assign_code(const std::string& a) { return a == "NGP" ? HM_NGP : (a == "CIC" ? HM_CIC : HM_TSC); }

struct Gpu {
    hm_ctx* ctx = nullptr;
    std::vector<void*> bufs;
    void* upload(const std::vector<double>& v) {
        void* d = hm_dev_alloc(ctx, sizeof(double) * v.size());
        if (!d) return nullptr;
        bufs.push_back(d);
        if (hm_memcpy_h2d(ctx, d, v.data(), sizeof(double) * v.size())) return nullptr;
        return d;
    }
    ~Gpu() {
        for (void* b : bufs) hm_dev_free(ctx, b);
        if (ctx) hm_ctx_destroy(ctx);
    }
};

// particle.hpp:377-420
static int print_survey_parameters(const Params& p, const SurveyCatalog& D, const SurveyCatalog& R,
                                   const hm_survey_info& info) {
    std::string path = p.output_dir + "/survey_parameters";
    FILE* fp = fopen(path.c_str(), "w");
    if (!fp) {
        printf("output directry '%s' does not exist\n", p.output_dir.c_str());
        return -1;
    }
    double dmin[3], dmax[3], rmin[3], rmax[3];
    D.min_max_shifted(info.shift, dmin, dmax);
    R.min_max_shifted(info.shift, rmin, rmax);
    fprintf(fp, "\n");
    fprintf(fp, "P_D.n_tot = %lld\n", D.n_tot);
    fprintf(fp, "P_D.n_tot_weight = %f\n", info.wsum_d);
    fprintf(fp, "P_R.n_tot = %lld\n", R.n_tot);
    fprintf(fp, "P_R.n_tot_weight = %f\n", info.wsum_r);
    fprintf(fp, "\n");
    fprintf(fp, "alpha = %.5f\n", info.alpha);
    fprintf(fp, "Vsurvey [(Gpc/h)^3] = %.5f\n", info.vsurvey / 1.0e9);
    fprintf(fp, "Vsurvey^(1/3) [Mpc/h] = %.5f\n", std::pow(info.vsurvey, 1.0 / 3.0));
    fprintf(fp, "nbar / 1.0e-4 [(Mpc/h)^{-3}] = %.5f\n", info.nbar / 1.0e-4);
    fprintf(fp, "\n");
    for (int a = 0; a < 3; a++) fprintf(fp, "P_D.pos_min[%d] [Mpc/h] = %.5f\n", a, dmin[a]);
    for (int a = 0; a < 3; a++) fprintf(fp, "P_R.pos_min[%d] [Mpc/h] = %.5f\n", a, rmin[a]);
    fprintf(fp, "\n");
    for (int a = 0; a < 3; a++) fprintf(fp, "P_D.pos_max[%d] [Mpc/h] = %.5f\n", a, dmax[a]);
    for (int a = 0; a < 3; a++) fprintf(fp, "P_R.pos_max[%d] [Mpc/h] = %.5f\n", a, rmax[a]);
    fprintf(fp, "\n");
    fclose(fp);
    return 0;
}

int main(int argc, char* argv[]) {
    g_t0 = now();
    printf("\n");
    printf("Start: %.3f sec.\n", 0.0);
    printf("\n");
    stage("Reading input parameters");
    if (argc != 2) {
        printf("The input parameter file is not found.\n");
        printf("For example, run './a.out param.ini'\n");
        EXIT();
    }
    Params p;
    if (p.read(argv[1])) {
        printf("'param.readParameters' fails.\n");
        EXIT();
    }
    if (p.make_output_directory()) {
        printf("'param.makeOutputDirectory' fails.\n");
        EXIT();
    }
    if (p.print_parameters()) {
        printf("'param.printParameters' fails.\n");
        EXIT();
    }
    std::vector<double> kbin(p.n_kbin), rbin(p.n_rbin);
    set_kbin(p, kbin.data());
    set_rbin(p, rbin.data());
    stage("Done");

    printf("\n");
    stage("Reading particle data");
    SurveyCatalog D, R;
    if (D.read(p.data_file)) {
        printf("'P_D.readParticleBOSS' fails.\n");
        EXIT();
    }
    if (R.read(p.random_file)) {
        printf("'P_R.readParticleBOSS' fails.\n");
        EXIT();
    }
    const bool recon = p.flag_recon == "True";
    int measure = -1;
    std::string fails;
    if (p.measure == "pk") {
        measure = HM_SURVEY_PK;
        fails = "'calcPowerSpectrum(P_D, P_R, param, alpha, kbin, Vsurvey)' fails.";
    } else if (p.measure == "2PCF") {
        measure = HM_SURVEY_XI;
        fails = "'calcTwoPointFunction(P_D, P_R, param, alpha, rbin, Vsurvey)' fails.";
    } else if (p.measure == "window2PCF") {
        measure = HM_SURVEY_XI_WINDOW;
        fails = "'calcTwoPointWindowCorrelationFunction(P_R, param, rbin, Vsurvey)' fails.";
    } else if (p.measure == "bk") {
        measure = HM_SURVEY_BK;
        fails = "'calcBiSpectrum(P_D, P_R, param, alpha, kbin, Vsurvey)' fails.";
    } else if (p.measure == "3PCF") {
        measure = HM_SURVEY_ZETA;
        fails = "'calcThreePointCorrelationFunction(P_D, P_R, param, alpha, kbin, Vsurvey)' fails.";
    } else if (p.measure == "window3PCF") {
        measure = HM_SURVEY_ZETA_WINDOW;
        fails = "'calcThreePointWindowCorrelationFunction(P_R, param, rbin, Vsurvey)' fails.";
    } else {
        printf("measure = %s is not available in this build.\n", p.measure.c_str());
        EXIT();
    }
    const bool three = measure >= HM_SURVEY_BK;
    // bispec.hpp:12-18, 363-369, 624-630: wigner_3j(ell1, ell2, ELL, 0, 0, 0) vanishes unless
    // ell1 + ell2 + ELL is even and |ell1 - ell2| <= ELL <= ell1 + ell2
    if (three && (((p.ell1 + p.ell2 + p.ELL) & 1) || p.ELL < std::abs(p.ell1 - p.ell2) || p.ELL > p.ell1 + p.ell2)) {
        printf("This multipole combination is not allowed.\n");
        printf("It should be wigner_3j(ell1, ell2, ELL, 0,0,0) != 0\n");
        printf("%s\n", fails.c_str());
        EXIT();
    }
    // powerspec.hpp:11-15, 101-105, 188-192
    if (!three && !((p.ELL == p.ell1) && (p.ell2 == 0))) {
        printf("This multipole combination is not allowed.\n");
        printf("It should be ELL == ell1 and ell2 == 0\n");
        printf("%s\n", fails.c_str());
        EXIT();
    }
    Gpu g;
    g.ctx = hm_ctx_create(0, 0);
    if (!g.ctx) {
        printf("no usable GPU (hm_ctx_create failed)\n");
        EXIT();
    }
    void* dpD = g.upload(D.pos);
    void* dwD = g.upload(D.w);
    void* dpR = g.upload(R.pos);
    void* dwR = g.upload(R.w);
    if (!dpD || !dwD || !dpR || !dwR) {
        printf("particle upload failed: %s\n", hm_last_error(g.ctx));
        EXIT();
    }
    hm_survey_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    for (int a = 0; a < 3; a++) {
        cfg.mesh.n_mesh[a] = p.n_mesh[a];
        cfg.mesh.boxsize[a] = p.boxsize[a];
    }
    cfg.mesh.assign = assign_code(p.assign);
    cfg.precision = p.precision == "FP64" ? HM_FP64 : HM_FP32;
    cfg.n_kbin = p.n_kbin;
    cfg.kmin = p.kmin;
    cfg.kmax = p.kmax;
    cfg.ell = p.ELL;
    cfg.prune = 1;
    cfg.measure = measure;
    cfg.n_rbin = p.n_rbin;
    cfg.rmin = p.rmin;
    cfg.rmax = p.rmax;
    cfg.recon = recon ? 1 : 0;
    for (int a = 0; a < 3; a++) cfg.n_mesh_recon[a] = p.n_mesh_recon[a];
    cfg.b1 = p.b1_fid;
    cfg.RG = p.RG;
    cfg.ell1 = p.ell1;
    cfg.ell2 = p.ell2;
    hm_survey_plan* plan = hm_survey_plan_create(g.ctx, &cfg);
    if (!plan) {
        printf("hm_survey_plan_create: %s\n", hm_last_error(g.ctx));
        EXIT();
    }
    const int nbin = (measure == HM_SURVEY_PK || measure == HM_SURVEY_BK) ? p.n_kbin : p.n_rbin;
    std::vector<double> pk(2 * (size_t)nbin * (three ? nbin : 1));
    std::vector<long long> nm(nbin);
    hm_survey_info info;
    printf("Calculating '%s': %.3f sec.\n", p.measure.c_str(), now() - g_t0);
    if (hm_survey_plan_execute(g.ctx, plan, (const double*)dpD, (const double*)dwD, D.n_tot, (const double*)dpR,
                               (const double*)dwR, R.n_tot, &info, pk.data(), nm.data())) {
        printf("hm_survey_plan_execute: %s\n", hm_last_error(g.ctx));
        printf("%s\n", fails.c_str());
        hm_survey_plan_destroy(g.ctx, plan);
        EXIT();
    }
    hm_survey_plan_destroy(g.ctx, plan);
    if (print_survey_parameters(p, D, R, info)) {
        printf("'ParticleBOSSClass::printParameters(param, P_D, P_R, alpha, Vsurvey, nbar))' fails.\n");
        EXIT();
    }
    char buf[2048];
    if (measure == HM_SURVEY_PK) {
        // powerspec.hpp:72-87
        snprintf(buf, sizeof(buf), recon ? "%s/pk%d_recon.dat" : "%s/pk%d.dat", p.output_dir.c_str(), p.ELL);
        FILE* fp = fopen(buf, "w");
        if (!fp) EXIT();
        fprintf(fp, "## k [h/Mpc]\t pk_real [(Mpc/h)^3]\t pk_imag [(Mpc/h)^3]\n");
        for (int i = 0; i < p.n_kbin; i++) fprintf(fp, "%.5f \t %.7e \t %.7e\n", kbin[i], pk[2 * i], pk[2 * i + 1]);
        fclose(fp);
    } else if (measure == HM_SURVEY_BK) {
        // bispec.hpp:329-346, one file per ith_kbin (all rows with bk_all_rows = True)
        const int r0 = p.bk_all_rows == "True" ? 0 : p.ith_kbin;
        const int r1 = p.bk_all_rows == "True" ? nbin : p.ith_kbin + 1;
        for (int ith = r0; ith < r1; ith++) {
            snprintf(buf, sizeof(buf), recon ? "%s/bk%d%d%d_%02d_recon.dat" : "%s/bk%d%d%d_%02d.dat",
                     p.output_dir.c_str(), p.ell1, p.ell2, p.ELL, ith);
            FILE* fp = fopen(buf, "w");
            if (!fp) EXIT();
            fprintf(fp, "## k1 [h/Mpc] \t k2 [h/Mpc] \t bk_real \t bk_imag\n");
            for (int i = 0; i < nbin; i++) {
                const size_t o = 2 * ((size_t)ith * nbin + i);
                fprintf(fp, "%.5f \t %.5f \t %.7e \t %.7e\n", kbin[ith], kbin[i], pk[o], pk[o + 1]);
            }
            fclose(fp);
        }
    } else if (three) {
        // bispec.hpp:591-606, 851-866: real part, one file per ith_rbin
        const bool win = measure == HM_SURVEY_ZETA_WINDOW;
        const int r0 = p.bk_all_rows == "True" ? 0 : p.ith_rbin;
        const int r1 = p.bk_all_rows == "True" ? nbin : p.ith_rbin + 1;
        for (int ith = r0; ith < r1; ith++) {
            const char* fmt = win ? (recon ? "%s/zeta%d%d%d_%02d_recon_window.dat" : "%s/zeta%d%d%d_%02d_window.dat")
                                  : (recon ? "%s/zeta%d%d%d_%02d_recon.dat" : "%s/zeta%d%d%d_%02d.dat");
            snprintf(buf, sizeof(buf), fmt, p.output_dir.c_str(), p.ell1, p.ell2, p.ELL, ith);
            FILE* fp = fopen(buf, "w");
            if (!fp) EXIT();
            fprintf(fp, "## r1 [Mpc/h] \t r2 [Mpc/h] \t zeta \n");
            for (int i = 0; i < nbin; i++)
                fprintf(fp, "%.5f \t %.5f \t %.7e\n", rbin[ith], rbin[i], pk[2 * ((size_t)ith * nbin + i)]);
            fclose(fp);
        }
    } else {
        // powerspec.hpp:161-175, 253-266
        const char* fmt = measure == HM_SURVEY_XI ? (recon ? "%s/xi%d_recon.dat" : "%s/xi%d.dat")
                                                  : (recon ? "%s/xi%d_recon_window.dat" : "%s/xi%d_window.dat");
        snprintf(buf, sizeof(buf), fmt, p.output_dir.c_str(), p.ELL);
        FILE* fp = fopen(buf, "w");
        if (!fp) EXIT();
        fprintf(fp, "## r [Mpc/h] \t xi\n");
        for (int i = 0; i < p.n_rbin; i++) fprintf(fp, "%.5f \t %.7e\n", rbin[i], pk[2 * i]);
        fclose(fp);
    }
    stage("Done");
    printf("\n");
    printf("End: %.3f sec.\n", now() - g_t0);
    return 0;
}
