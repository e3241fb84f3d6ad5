// hitomi_box -- drop-in for the periodic-box program of hitomi_measurement
// (main_box.cpp:13-242): `hitomi_box param.ini`.
//
// Reads the same .ini keys, the same Gadget-1 / Rockstar inputs, writes the
// same input_parameters, simulation_parameters and pk%d.dat /
// bk%d%d%d_%02d.dat files, and exits with status 1 where the reference calls
// EXIT().  The measurement runs on the GPU through libhitomi_amd.so.
//
// Supported: measure = pk and 2PCF (any ELL, ell1 = ELL, ell2 = 0) and
// measure = bk and 3PCF for any (ell1, ell2, ELL) with
// wigner_3j(ell1, ell2, ELL, 0, 0, 0) != 0, each with or without
// reconstruction (flag_recon).
// Other measurements are rejected with an error instead of being computed
// differently.
#include "../../include/hitomi_amd.h"
#include "hm_catalog.hpp"
#include "hm_param.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
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
    double* d_pos = nullptr;
    double* d_pos_r = nullptr;  // randoms of a reconstructed box
    long long n_r = 0;
    double alpha = 0.0;
    ~Gpu() {
        if (ctx && d_pos) hm_dev_free(ctx, d_pos);
        if (ctx && d_pos_r) hm_dev_free(ctx, d_pos_r);
        if (ctx) hm_ctx_destroy(ctx);
    }
};

static double* upload(Gpu& g, const std::vector<double>& v) {
    double* d = (double*)hm_dev_alloc(g.ctx, sizeof(double) * (v.empty() ? 1 : v.size()));
    if (!d) return nullptr;
    if (!v.empty() && hm_memcpy_h2d(g.ctx, d, v.data(), sizeof(double) * v.size())) return nullptr;
    return d;
}

enum Measure { M_PK, M_XI, M_BK, M_ZETA };

static int run_box(const Params& p, const Catalog& P_D, Gpu& g, Measure what, int ell, std::vector<double>& pk,
                   std::vector<double>& bk, std::vector<double>& xi, std::vector<double>* zeta = nullptr) {
    const bool want_bk = what == M_BK;
    hm_box_config cfg;
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
    cfg.n_pk_ells = what == M_PK ? 1 : 0;
    cfg.pk_ells[0] = ell;
    cfg.bk_ell[0] = p.ell1;
    cfg.bk_ell[1] = p.ell2;
    cfg.bk_ell[2] = p.ELL;
    cfg.bk = want_bk ? 1 : 0;
    cfg.regrid = 1;
    cfg.prune = 1;
    cfg.n_rbin = p.n_rbin;
    cfg.rmin = p.rmin;
    cfg.rmax = p.rmax;
    cfg.n_xi_ells = what == M_XI ? 1 : 0;
    cfg.xi_ells[0] = ell;
    if (what == M_ZETA) {
        cfg.zeta = 1;
        cfg.zeta_ell[0] = p.ell1;
        cfg.zeta_ell[1] = p.ell2;
        cfg.zeta_ell[2] = p.ELL;
    }
    hm_box_plan* plan = hm_box_plan_create(g.ctx, &cfg);
    if (!plan) {
        printf("hm_box_plan_create: %s\n", hm_last_error(g.ctx));
        return -1;
    }
    const int nk = p.n_kbin;
    pk.assign(2 * (size_t)nk, 0.0);
    bk.assign(2 * (size_t)nk * nk, 0.0);
    xi.assign((size_t)p.n_rbin, 0.0);
    hm_box_result res;
    std::memset(&res, 0, sizeof(res));
    if (what == M_PK) res.pk = pk.data();
    if (want_bk) res.bk = bk.data();
    if (what == M_XI) res.xi = xi.data();
    if (what == M_ZETA && zeta) {
        zeta->assign(2 * (size_t)p.n_rbin * p.n_rbin, 0.0);
        res.zeta = zeta->data();
    }
    int rc = g.d_pos_r ? hm_box_plan_execute_recon(g.ctx, plan, g.d_pos, P_D.n_tot, g.d_pos_r, g.n_r, g.alpha, &res)
                       : hm_box_plan_execute(g.ctx, plan, g.d_pos, P_D.n_tot, &res);
    if (rc) printf("hm_box_plan_execute: %s\n", hm_last_error(g.ctx));
    hm_box_plan_destroy(g.ctx, plan);
    return rc;
}

int main(int argc, char* argv[]) {
    g_t0 = now();
    printf("Start: %.3f sec.\n", 0.0);
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

    stage("Reading particle data");
    Catalog P_D;
    if (p.sim_data == "Gadget") {
        if (P_D.read_gadget(p.data_file)) {
            printf("'P_D.readParticlesFromGadgetSnapshot(param.data_file)' fails.\n");
            EXIT();
        }
    } else if (p.sim_data == "Rockstar") {
        if (P_D.read_rockstar(p.data_file, p.log10_Mmin, p.log10_Mmax)) {
            printf("'P_D.readParticlesFromRockstar(param.data_file, param.log10_Mmin, param.log10_Mmax)' fails.\n");
            EXIT();
        }
    } else {
        EXIT();
    }
    // particleGadget.hpp:643-653: the box size comes from the catalogue
    for (int a = 0; a < 3; a++) p.boxsize[a] = P_D.boxsize;
    p.volume = p.boxsize[0] * p.boxsize[1] * p.boxsize[2];
    if (p.flag_RSD == "True") {
        if (P_D.set_rsd(p.boxsize)) {
            printf("'P_D.setRSD(param)' fails.\n");
            EXIT();
        }
    }
    P_D.print_parameters(p.output_dir);
    stage("Done");

    Gpu g;
    g.ctx = hm_ctx_create(0, 0);
    if (!g.ctx) {
        printf("no usable GPU (hm_ctx_create failed)\n");
        EXIT();
    }
    g.d_pos = upload(g, P_D.pos);
    if (!g.d_pos) {
        printf("particle upload failed: %s\n", hm_last_error(g.ctx));
        EXIT();
    }
    const bool recon = p.flag_recon == "True";
    const char* sfx = recon ? "_recon" : "";
    if (recon) {
        // main_box.cpp:141-159
        printf("Calculating reconstruction: %.3f sec.\n", now() - g_t0);
        Catalog P_R;
        if (P_R.read_randoms(p.random_file)) {
            // the reference prints this and carries on with no randoms (then divides by N_R = 0)
            printf("'P_R.readRandomParticles(param.random_file)' fails.\n");
            EXIT();
        }
        g.d_pos_r = upload(g, P_R.pos);
        g.n_r = P_R.n_tot;
        if (!g.d_pos_r) {
            printf("particle upload failed: %s\n", hm_last_error(g.ctx));
            EXIT();
        }
        hm_recon_config rcfg;
        std::memset(&rcfg, 0, sizeof(rcfg));
        for (int a = 0; a < 3; a++) {
            rcfg.mesh.n_mesh[a] = p.n_mesh_recon[a];
            rcfg.mesh.boxsize[a] = p.boxsize[a];
        }
        rcfg.mesh.assign = assign_code(p.assign);
        rcfg.precision = p.precision == "FP64" ? HM_FP64 : HM_FP32;
        rcfg.b1 = p.b1_fid;
        rcfg.RG = p.RG;
        // particle_reconstruction.hpp:221-222: alpha by integer division (Q4), Vsurvey = V
        const double alpha_int = double(P_D.n_tot / P_R.n_tot);
        if (hm_recon_displace(g.ctx, &rcfg, g.d_pos, nullptr, P_D.n_tot, g.d_pos_r, nullptr, P_R.n_tot, alpha_int,
                              p.volume / double(P_D.n_tot)) ||
            hm_periodic_wrap(g.ctx, g.d_pos, P_D.n_tot, p.boxsize) ||
            hm_periodic_wrap(g.ctx, g.d_pos_r, P_R.n_tot, p.boxsize) ||
            hm_memcpy_d2h(g.ctx, P_D.pos.data(), g.d_pos, sizeof(double) * 3 * (size_t)P_D.n_tot) ||
            hm_memcpy_d2h(g.ctx, P_R.pos.data(), g.d_pos_r, sizeof(double) * 3 * (size_t)P_R.n_tot)) {
            printf("reconstruction failed: %s\n", hm_last_error(g.ctx));
            EXIT();
        }
        P_D.calc_min_max();
        P_R.calc_min_max();
        g.alpha = double(P_D.n_tot) / double(P_R.n_tot);  // main_box.cpp:153
        if (Catalog::print_parameters_recon(p.output_dir, P_D, P_R, g.alpha)) EXIT();
        stage("Done");
    }

    printf("Calculating '%s': %.3f sec.\n", p.measure.c_str(), now() - g_t0);
    const int nk = p.n_kbin;
    std::vector<double> pk, bk, xi;
    if (p.measure == "pk") {
        // powerspec.hpp:280-284
        if (!((p.ELL == p.ell1) && (p.ell2 == 0))) {
            printf("This multipole combination is not allowed.\n");
            printf("It should be ELL == ell1 and ell2 == 0\n");
            printf("'calcPowerSpectrumForBOX(P_D, param, kbin)' fails.\n");
            EXIT();
        }
        if (run_box(p, P_D, g, M_PK, p.ELL, pk, bk, xi)) {
            printf("'calcPowerSpectrumForBOX(P_D, param, kbin)' fails.\n");
            EXIT();
        }
        // powerspec.hpp:318-327
        char buf[2048];
        snprintf(buf, sizeof(buf), "%s/pk%d%s.dat", p.output_dir.c_str(), p.ELL, sfx);
        FILE* fp = fopen(buf, "w");
        if (!fp) EXIT();
        fprintf(fp, "## k [h/Mpc]\t pk_real [(Mpc/h)^3]\t pk_imag [(Mpc/h)^3]\n");
        for (int i = 0; i < nk; i++) fprintf(fp, "%.5f \t %.7e \t %.7e\n", kbin[i], pk[2 * i], pk[2 * i + 1]);
        fclose(fp);
    } else if (p.measure == "2PCF") {
        // powerspec.hpp:341-345
        if (!((p.ELL == p.ell1) && (p.ell2 == 0))) {
            printf("This multipole combination is not allowed.\n");
            printf("It should be ELL == ell1 and ell2 == 0\n");
            printf("'calcTwoPointCorrelationFunctionForBOX(P_D, param, rbin)' fails.\n");
            EXIT();
        }
        if (run_box(p, P_D, g, M_XI, p.ELL, pk, bk, xi)) {
            printf("'calcTwoPointCorrelationFunctionForBOX(P_D, param, rbin)' fails.\n");
            EXIT();
        }
        // powerspec.hpp:380-386
        char buf[2048];
        snprintf(buf, sizeof(buf), "%s/xi%d%s.dat", p.output_dir.c_str(), p.ELL, sfx);
        FILE* fp = fopen(buf, "w");
        if (!fp) EXIT();
        fprintf(fp, "## r [Mpc/h] \t xi\n");
        for (int i = 0; i < p.n_rbin; i++) fprintf(fp, "%.5f \t %.7e\n", rbin[i], xi[i]);
        fclose(fp);
    } else if (p.measure == "bk") {
        if (run_box(p, P_D, g, M_BK, 0, pk, bk, xi)) {
            printf(recon ? "'calcBiSpectrumForBOXForReconstruction(P_D, P_R, param, alpha, kbin)' fails.\n"
                         : "'calcBiSpectrumForBOX(P_D, param, kbin)' fails.\n");
            EXIT();
        }
        // bispec.hpp:1187-1196, one file per ith_kbin (all rows with bk_all_rows = True)
        const int r0 = p.bk_all_rows == "True" ? 0 : p.ith_kbin;
        const int r1 = p.bk_all_rows == "True" ? nk : p.ith_kbin + 1;
        for (int ith = r0; ith < r1; ith++) {
            char buf[2048];
            snprintf(buf, sizeof(buf), "%s/bk%d%d%d_%02d%s.dat", p.output_dir.c_str(), p.ell1, p.ell2, p.ELL, ith, sfx);
            FILE* fp = fopen(buf, "w");
            if (!fp) EXIT();
            fprintf(fp, "## k1 [h/Mpc] \t k2 [h/Mpc] \t bk_real \t bk_imag\n");
            for (int i = 0; i < nk; i++) {
                const size_t o = 2 * ((size_t)ith * nk + i);
                fprintf(fp, "%.5f \t %.5f \t %.7e \t %.7e\n", kbin[ith], kbin[i], bk[o], bk[o + 1]);
            }
            fclose(fp);
        }
    } else if (p.measure == "3PCF") {
        std::vector<double> zeta;
        if (run_box(p, P_D, g, M_ZETA, 0, pk, bk, xi, &zeta)) {
            printf(recon ? "'calcThreePointCorrelationFunctionForBOXForReconstruction(P_D, P_R, param, alpha, rbin)' fails.\n"
                         : "'calcThreePointCorrelationFunctionForBOX(P_D, param, rbin)' fails.\n");
            EXIT();
        }
        // bispec.hpp:1409-1417, one file per ith_rbin (all rows with bk_all_rows = True)
        const int nr = p.n_rbin;
        const int r0 = p.bk_all_rows == "True" ? 0 : p.ith_rbin;
        const int r1 = p.bk_all_rows == "True" ? nr : p.ith_rbin + 1;
        for (int ith = r0; ith < r1; ith++) {
            char buf[2048];
            snprintf(buf, sizeof(buf), "%s/zeta%d%d%d_%02d%s.dat", p.output_dir.c_str(), p.ell1, p.ell2, p.ELL, ith,
                     sfx);
            FILE* fp = fopen(buf, "w");
            if (!fp) EXIT();
            fprintf(fp, "## r1 [Mpc/h] \t r2 [Mpc/h] \t zeta \n");
            for (int i = 0; i < nr; i++)
                fprintf(fp, "%.5f \t %.5f \t %.7e\n", rbin[ith], rbin[i], zeta[2 * ((size_t)ith * nr + i)]);
            fclose(fp);
        }
    } else {
        printf("measure = %s is not available in this build.\n", p.measure.c_str());
        EXIT();
    }
    stage("Done");
    printf("End: %.3f sec.\n", now() - g_t0);
    return 0;
}
