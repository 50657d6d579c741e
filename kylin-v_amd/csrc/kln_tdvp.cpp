// kln-tdvp: the TDVP command-line driver on the MI355X engine, through the C ABI.
//
// Same flags and file contract as the reference's src/kln-tdvp.cpp:10-80:
//   -w MPS file, -o MPO file, -s "[nK,n2,nSA,n1]" stage list (holstein/tdvp.h:19-32),
//   -v time step, -D max bond dimension (default 1000), -t SVD cutoff (default 1e-8);
//   Krylov tolerance 1e-8 and 15 iterations fixed (src/kln-tdvp.cpp:27-28).
// After every sweep the MPS is written to State/T-<sweep>.mps in the reference's binary
// format (holstein/tdvp.h:325-338) and "Max bond dimension = D on bond i-j" is printed
// (holstein/state.h:187-200).  Extra flags: -p fp64|fp32|bf16x3 (H_eff contraction precision),
// -g device, -r seed of the SA-1TDVP draws (default 1; sweep k draws under seed + k).
// The nK stage runs global Krylov steps (holstein/krylov.h:206-284) before the sweeps, as
// holstein/tdvp.h:318-327.  Every two-site SVD appends to ./SVS (holstein/tdvp.h:71-74,
// holstein/svd.h:390-400); the nSA stage fits the per-bond widths to it (bin/nls.py, run
// in-process), writes them to ./beta, reads them back as betaSet = 1 / width
// (holstein/tdvp.h:341-358) and runs the SA-1TDVP sweeps (holstein/tdvp.h:359-373).
#include <getopt.h>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kylinv.h"

static void die(kv_ctx* ctx, int rc, const char* what) {
    std::fprintf(stderr, "kln-tdvp: %s failed (%d): %s\n", what, rc, ctx ? kv_last_error(ctx) : "");
    std::exit(1);
}

static bool parse_stage(const std::string& s, int st[4]) {
    // drop the first character, then four comma-separated integers
    if (s.size() < 2) return false;
    std::string body = s.substr(1);
    for (int i = 0; i < 4; ++i) {
        size_t pos = body.find(',');
        std::string tok = body.substr(0, pos);
        st[i] = std::atoi(tok.c_str());
        if (pos == std::string::npos) {
            if (i != 3) return false;
            break;
        }
        body = body.substr(pos + 1);
    }
    return true;
}

int main(int argc, char** argv) {
    static const option long_op[] = {{"wavefunc", required_argument, nullptr, 'w'},
                                     {"operator", required_argument, nullptr, 'o'},
                                     {"stage", required_argument, nullptr, 's'},
                                     {"interval", required_argument, nullptr, 'v'},
                                     {"max-bond", required_argument, nullptr, 'D'},
                                     {"tolerance", required_argument, nullptr, 't'},
                                     {"precision", required_argument, nullptr, 'p'},
                                     {"device", required_argument, nullptr, 'g'},
                                     {"seed", required_argument, nullptr, 'r'},
                                     {"help", no_argument, nullptr, 'h'},
                                     {nullptr, 0, nullptr, 0}};
    std::string mps_file, mpo_file, stage = "[0,0,0,0]";
    double dt = 0.0, tol = 1e-8;
    int maxdim = 1000, prec = KV_FP64, device = 0, ch, idx = 0;
    unsigned long long seed = 1;
    while ((ch = getopt_long(argc, argv, "w:o:s:v:D:t:p:g:r:h", long_op, &idx)) != -1) {
        switch (ch) {
            case 'w': mps_file = optarg; break;
            case 'o': mpo_file = optarg; break;
            case 's': stage = optarg; break;
            case 'v': dt = std::atof(optarg); break;
            case 'D': maxdim = std::atoi(optarg); break;
            case 't': tol = std::atof(optarg); break;
            case 'p':
                prec = std::strcmp(optarg, "fp32") == 0     ? KV_FP32
                       : std::strcmp(optarg, "bf16x3") == 0 ? KV_BF16X3
                                                            : KV_FP64;
                break;
            case 'g': device = std::atoi(optarg); break;
            case 'r': seed = std::strtoull(optarg, nullptr, 10); break;
            case 'h':
                std::printf("Program to calculate tDMRG (TDVP) on an MI355X\n"
                            "Usage: kln-tdvp -w [MPS file] -o [MPO file] -s [n_K,n_2,n_sa,n_1] "
                            "-v [dt] -D [max bond] -t [SVD cutoff] [-p fp64|fp32|bf16x3] [-g device] "
                            "[-r seed]\n");
                return 0;
            default: std::printf("Unknown Option!\n"); break;
        }
    }
    int st[4];
    if (!parse_stage(stage, st)) {
        std::fprintf(stderr, "kln-tdvp: bad stage string %s\n", stage.c_str());
        return 1;
    }
    kv_ctx* ctx = nullptr;
    int rc = kv_ctx_create(device, &ctx);
    if (rc) die(nullptr, rc, "kv_ctx_create");
    int ns = 0;
    if ((rc = kv_mps_load(ctx, mps_file.c_str(), 0, &ns, nullptr))) die(ctx, rc, "reading the MPS");
    std::vector<kv_tensor*> mps(ns), mpo(ns);
    if ((rc = kv_mps_load(ctx, mps_file.c_str(), ns, &ns, mps.data()))) die(ctx, rc, "reading the MPS");
    int nso = 0;
    if ((rc = kv_mpo_load(ctx, mpo_file.c_str(), ns, &nso, mpo.data()))) die(ctx, rc, "reading the MPO");
    if (nso != ns) {
        std::fprintf(stderr, "kln-tdvp: MPS has %d sites, MPO %d\n", ns, nso);
        return 1;
    }
    kv_tdvp_config cfg{dt, tol, maxdim, 1e-8, 15, prec};
    std::printf("Krylov iteration threshold = %g\nKrylov maximal iteration   = 15\n"
                "Time step                  = %g\nSVD threshold              = %g\n"
                "SVD maximal allowed        = %d\n", 1e-8, dt, tol, maxdim);
    kv_tdvp* drv = nullptr;
    if ((rc = kv_tdvp_create(ctx, ns, mps.data(), mpo.data(), &cfg, &drv))) die(ctx, rc, "TDVP setup");
    if ((rc = kv_tdvp_svs_open(drv, "SVS"))) die(ctx, rc, "opening the SVS log");
    mkdir("State", 0755);
    std::vector<int> bonds(ns > 1 ? ns - 1 : 1);
    std::vector<double> betas;
    int swp = 0;
    const char* stage_name[4] = {"Krylov", "2TDVP", "SA-1TDVP", "1TDVP"};
    for (int stage_i = 0; stage_i < 4; ++stage_i) {
        const int n = st[stage_i];
        if (stage_i == 2) {
            if (n == 0) continue;
            int nb = 0;
            if ((rc = kv_sa_fit_widths("SVS", nullptr, 0, &nb))) die(ctx, rc, "fitting the SVS widths");
            std::vector<double> w(nb > 0 ? nb : 1);
            if ((rc = kv_sa_fit_widths("SVS", w.data(), nb, &nb))) die(ctx, rc, "fitting the SVS widths");
            FILE* fb = std::fopen("beta", "w");
            if (!fb) die(ctx, KV_ERR_ARG, "writing the beta file");
            for (int b = 0; b < nb; ++b) std::fprintf(fb, "%21.15f\n", w[b]);
            std::fclose(fb);
            std::printf("Fitting OK!\n");
            fb = std::fopen("beta", "r");
            if (!fb) die(ctx, KV_ERR_ARG, "reading the beta file");
            char line[256];
            while (std::fgets(line, sizeof line, fb))
                if (std::strlen(line) > 4) betas.push_back(1.0 / std::atof(line));
            std::fclose(fb);
            std::printf("Read betas OK!\n");
            if ((int)betas.size() < ns - 1) {
                std::fprintf(stderr, "kln-tdvp: the SA-1TDVP stage needs a width for each of the %d bonds "
                             "(%d fitted: run two-site sweeps first)\n", ns - 1, (int)betas.size());
                return 1;
            }
        }
        std::printf("\n ---------------- Start %s stage ----------------\n", stage_name[stage_i]);
        for (int k = 0; k < n; ++k, ++swp) {
            if (stage_i > 0) std::printf("\nSweep %d:\n", swp + 1);
            rc = stage_i == 0   ? kv_tdvp_global_krylov(drv, nullptr)
                 : stage_i == 1 ? kv_tdvp_two_site(drv)
                 : stage_i == 2 ? kv_tdvp_sa_one_site(drv, betas.data(), (int)betas.size(), seed + swp)
                                : kv_tdvp_one_site(drv);
            if (rc) die(ctx, rc, "sweep");
            kv_tdvp_bond_dims(drv, bonds.data());
            int mb = 0, mbp = 0;
            for (int i = 0; i + 1 < ns; ++i)
                if (bonds[i] > mbp) {
                    mbp = bonds[i];
                    mb = i;
                }
            std::printf("Max bond dimension = %d on bond %d-%d\n", mbp, mb, mb + 1);
            std::vector<kv_tensor*> sites(ns);
            for (int i = 0; i < ns; ++i)
                if ((rc = kv_tdvp_get_site(drv, i, &sites[i]))) die(ctx, rc, "reading a site");
            const std::string fn = "State/T-" + std::to_string(swp) + ".mps";
            if ((rc = kv_mps_save(ctx, fn.c_str(), ns, sites.data()))) die(ctx, rc, "writing the MPS");
            for (auto* t : sites) kv_tensor_destroy(t);
        }
    }
    double s[22] = {0};
    kv_tdvp_stats(drv, s, 22);
    const char* names[6] = {"Evl-2", "Evl-1", "Evl-0", "SVD", "Update", "Sweep"};
    std::printf("\n%-12s %12s\n", "Module", "Wall(s)");
    for (int i = 0; i < 6; ++i) std::printf("%-12s %12.4f\n", names[i], s[i]);
    std::printf("%-12s %12.4f  (%.3g flop, %.1f TFLOP/s)\n", "H_eff GPU", s[6], s[7],
                s[6] > 0 ? s[7] / s[6] / 1e12 : 0.0);
    kv_tdvp_destroy(drv);
    for (int i = 0; i < ns; ++i) {
        kv_tensor_destroy(mps[i]);
        kv_tensor_destroy(mpo[i]);
    }
    kv_ctx_destroy(ctx);
    return 0;
}
