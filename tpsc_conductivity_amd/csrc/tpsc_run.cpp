// tpsc_run -- input.dat driver over the C ABI (tpsc_main.cpp:25-114 and :243).
//
//   tpsc_run [input.dat] [--out DIR] [--mixed] [--no-chijj]
//
// Reads the reference's input file, one line "U tp tpp t3 n T nk0 Nw0 chijj3 mmin"
// (tpsc_main.cpp:33), echoes it the way the reference does, then runs the reference's
// sequence on the GPU: green(nk0, Nw0, n, T, eta = 0.1, {tp, tpp, t3, U}),
// calc_Self_FFT_spline, traceSelfG, find_mu (tpsc_main.cpp:61-75), then
// calc_chijj_vertex_corr_all_optim(chijj3, mmin, 1, NULL, 0) (tpsc_main.cpp:243), and writes
// the reference's text files into DIR (default: the current directory) with
// tpsc_write_files.  The post-processing of the reference's main (Fermi surfaces, MDC, EDC,
// Gwn_local) is not part of this driver.  Exit status 0 on success, 1 on a failed stage.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/tpsc_capi.h"

namespace {

int check(tpsc_ctx* c, int rc, const char* what)
{
    if (rc != TPSC_OK) std::cerr << what << " failed (" << rc << "): " << tpsc_last_error(c) << '\n';
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    std::string input = "input.dat", out;
    int precision = TPSC_PREC_FP64;
    bool chijj = true;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--out") && i + 1 < argc)
            out = argv[++i];
        else if (!strcmp(argv[i], "--mixed"))
            precision = TPSC_PREC_MIXED;
        else if (!strcmp(argv[i], "--no-chijj"))
            chijj = false;
        else if (argv[i][0] == '-') {
            std::cerr << "usage: tpsc_run [input.dat] [--out DIR] [--mixed] [--no-chijj]\n";
            return 2;
        } else
            input = argv[i];
    }
    double U, tp, tpp, t3, n, T;
    int nk0, Nw0, mmin;
    bool chijj3;
    std::ifstream in(input);
    if (!in) {
        std::cout << "echec d'ouverture de input.dat" << '\n';
        return 1;
    }
    in >> U >> tp >> tpp >> t3 >> n >> T >> nk0 >> Nw0 >> chijj3 >> mmin;
    if (!in) {
        std::cerr << input << ": expected U tp tpp t3 n T nk0 Nw0 chijj3 mmin\n";
        return 1;
    }
    std::cout << "U: " << U << "  tp:  " << tp << "  tpp:  " << tpp << "  t3:  " << t3 << "  n:  " << n
              << "  T:  " << T << "  nk0:  " << nk0 << "  Nw0:  " << Nw0 << "  chijj3:  " << chijj3
              << "  mmin:  " << mmin << std::endl;

    tpsc_params p{};
    p.nk0 = nk0;
    p.Nw0 = Nw0;
    p.n = n;
    p.T = T;
    p.eta = 0.1;
    p.tp = tp;
    p.tpp = tpp;
    p.t3 = t3;
    p.U = U;
    p.coarse_pass = 1;
    p.precision = precision;
    tpsc_ctx* c = nullptr;
    if (check(c, tpsc_create(&p, &c), "tpsc_create")) {
        tpsc_destroy(c);
        return 1;
    }
    int rc = check(c, tpsc_calc_self_energy_full(c), "calc_Self_FFT_spline");
    double tr = 0, trc = 0, mu = 0;
    if (!rc) rc = check(c, tpsc_trace_self_G(c, &tr, &trc), "traceSelfG");
    if (!rc) rc = check(c, tpsc_find_mu(c, &mu), "find_mu");
    if (!rc && chijj) rc = check(c, tpsc_chijj(c, chijj3 ? 1 : 0, mmin, 1), "calc_chijj_vertex_corr_all_optim");
    double v = 0;
    tpsc_get_scalar(c, TPSC_S_USP, &v);
    std::cout << "Usp: " << v;
    tpsc_get_scalar(c, TPSC_S_UCH, &v);
    std::cout << "  Uch: " << v << "  mu: " << mu << '\n';
    int nfiles = 0;
    const int wrc = tpsc_write_files(c, out.empty() ? nullptr : out.c_str(), &nfiles);
    std::cout << nfiles << " files written" << (out.empty() ? "" : " to " + out) << '\n';
    tpsc_destroy(c);
    return (rc || wrc) ? 1 : 0;
}
