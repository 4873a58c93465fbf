// tpsc_files.cpp -- the reference's text outputs, written from a context through the public
// C ABI only (tpsc_get, tpsc_get_scalar, tpsc_get_params).
//
// File names use the reference's printf formats and the contents its iostream settings
// (precision, left / scientific flags, setw), so tools that read the reference's files read
// these.  Files the reference builds in several appends (TPSC_infos_*, TPSCparams_*) are
// written whole, with the lines in the reference's order.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/tpsc_capi.h"

namespace {

struct Writer {
    tpsc_ctx* c;
    std::string dir;
    int count = 0;
    int err = 0;

    std::string path(const char* fmt, ...) __attribute__((format(printf, 2, 3)))
    {
        char name[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(name, sizeof name, fmt, ap);
        va_end(ap);
        const char* base = (strncmp(name, "./", 2) == 0) ? name + 2 : name;
        return dir.empty() ? std::string(base) : dir + "/" + base;
    }
    bool open(std::ofstream& f, const std::string& p, std::ios::openmode m = std::ios::out)
    {
        f.open(p, m);
        if (!f) {
            err = TPSC_E_INVALID;
            return false;
        }
        count++;
        return true;
    }
    double scalar(int which)
    {
        double v = 0.0;
        tpsc_get_scalar(c, which, &v);
        return v;
    }
    std::vector<double> field(int f)
    {
        std::vector<double> v(tpsc_field_size(c, f) / sizeof(double));
        if (v.empty() || tpsc_get(c, f, v.data(), v.size() * sizeof(double))) v.clear();
        return v;
    }
};

}  // namespace

extern "C" int tpsc_write_files(tpsc_ctx* c, const char* dir, int* nfiles)
{
    if (!c) return TPSC_E_INVALID;
    tpsc_params p;
    tpsc_get_params(c, &p);
    Writer w{c, dir ? std::string(dir) : std::string()};
    const int stages = (int)w.scalar(TPSC_S_STAGES);
    const int nk = (int)w.scalar(TPSC_S_NK), nk0 = 2 * (nk - 1);
    const int nwb = (int)w.scalar(TPSC_S_NWB), nbw = 2 * (nwb - 1);
    const double U = p.U, tp = p.tp, tpp = p.tpp, t3 = p.t3, density = p.n, tem = p.T;

    // hamiltonien::calc_ek (hamiltonien.cpp:72-82)
    {
        const std::vector<double> ek = w.field(TPSC_F_EK);
        std::ofstream f;
        if (!ek.empty() && w.open(f, w.path("./e_k_%5.3f_%5.3f_%5.3f_%1d.dat", tp, tpp, t3, nk0))) {
            f << std::setiosflags(std::ios::left) << std::setprecision(15);
            for (int l = 0; l < nk; l++)
                for (int m = 0; m < nk; m++) f << std::setw(10) << l << std::setw(10) << m << ek[m + l * nk] << std::endl;
        }
    }
    // the green0 constructor's Fermi surfaces: find_FS(10001) (green0.cpp:626-676) and
    // find_FS_PBC (green0.cpp:51-105), at the finite-lattice mu0
    {
        const int nfs = 10001;
        std::vector<double> kxy(2 * (size_t)nfs);
        int cnt = 0;
        std::ofstream f;
        if (tpsc_fermi_surface(c, nfs, kxy.data(), &cnt) == TPSC_OK &&
            w.open(f, w.path("./FS_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d.dat", tp, tpp, t3, density, tem, nfs))) {
            f << std::setiosflags(std::ios::left) << std::setprecision(10);
            for (int j = 0; j < cnt; j++) f << std::setw(25) << kxy[2 * j] << kxy[2 * j + 1] << std::endl;
            f.close();
        }
        std::vector<int> mi(nk);
        std::vector<double> ak(nk);
        if (tpsc_fs_lattice(c, 0, mi.data(), ak.data()) == TPSC_OK) {
            const char* forms[2] = {"./FS_discret_AkF_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d.dat",
                                    "./FS_discret_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d.dat"};
            for (int k = 0; k < 2; k++) {
                if (!w.open(f, w.path(forms[k], tp, tpp, t3, density, tem, nk0))) continue;
                f << std::setiosflags(std::ios::left) << std::setprecision(8);
                for (int l = 0; l < nk; l++) {
                    f << std::setw(10) << l + mi[l] << std::setw(10) << mi[l];
                    if (k == 0) f << ak[l];
                    f << std::endl;
                }
                f.close();
            }
        }
    }
    // chi0::calc_chiqw_FFT_spline with save_chi0_q (chi0.cpp:720-737)
    if (stages & TPSC_STAGE_CHI0) {
        const std::vector<double> chi = w.field(TPSC_F_CHI0);
        const long nk8 = (long)nk * (nk + 1) / 2;
        const int l0 = (int)w.scalar(TPSC_S_QMAX_L), m0 = (int)w.scalar(TPSC_S_QMAX_M);
        std::ofstream f;
        if (!chi.empty() && w.open(f, w.path("./chi0_qmax_wn_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d_%d_%d.dat", tp, tpp,
                                             t3, density, tem, nk0, nbw, l0, m0))) {
            f << std::setiosflags(std::ios::left) << std::setprecision(16);
            for (int j = 0; j < nwb; j++) f << std::setw(10) << j << chi[m0 + (l0 * (l0 + 1)) / 2 + j * nk8] << '\n';
            f.close();
        }
        if (!chi.empty() &&
            w.open(f, w.path("./chi0_wn0_q_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat", tp, tpp, t3, density, tem, nk0, nbw))) {
            f << std::setiosflags(std::ios::left) << std::setprecision(16);
            for (int l = 0; l < nk; l++)
                for (int m = 0; m <= l; m++) f << std::setw(10) << l << std::setw(10) << m << chi[m + (l * (l + 1)) / 2] << '\n';
        }
    }
    // chi::find_Usp_Uch (chi.cpp:362-415), then traceSelfG (green.cpp:3088-3103), find_mu
    // (green.cpp:3259-3276) and traceSelfG2 (green.cpp:2978-2994)
    if (stages & TPSC_STAGE_USP_UCH) {
        const double Usp = w.scalar(TPSC_S_USP), Uch = w.scalar(TPSC_S_UCH);
        const double difRelsp = w.scalar(TPSC_S_DIFREL_SP), difRelch = w.scalar(TPSC_S_DIFREL_CH);
        const double chi0max = w.scalar(TPSC_S_CHI0MAX), dblOcc = w.scalar(TPSC_S_DBLOCC);
        const int q0 = (int)w.scalar(TPSC_S_QMAX_L), q1 = (int)w.scalar(TPSC_S_QMAX_M);
        const double mu0 = w.scalar(TPSC_S_MU0), mu = w.scalar(TPSC_S_MU);
        const double chispmax = chi0max / (1.0 - Usp * chi0max / 2.0);
        std::ofstream f;
        if (w.open(f, w.path("./TPSC_infos_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat", U, tp, tpp, t3, density, tem,
                             nk0, nbw))) {
            f << std::setprecision(16);
            f << "Usp:  " << Usp << "   difRelsp:  " << difRelsp << '\n';
            f << "Uch:  " << Uch << "   difRelch:  " << difRelch << '\n';
            f << "chi0max: " << chi0max << "   chispmax: " << chispmax << "   rapport: " << chispmax / chi0max << '\n';
            f << "double occupation (ansatz): " << dblOcc << '\n';
            f << "qxmax, qymax:   " << q0 << ",  " << q1 << '\n';
            double xi[5];
            if ((stages & TPSC_STAGE_VERTEX) && tpsc_sp_corr_length(c, xi) == TPSC_OK) {
                // chi::sp_corr_length (chi.cpp:476-560), which calc_Self_FFT_spline runs before V
                const double vF = xi[0];
                f << "vitesse de Fermi aux points chauds:  " << vF << std::endl;
                const char* ax[2] = {"(axe y)", "(axe x)"};
                for (int a = 0; a < 2; a++) {
                    f << "longueur de correlation de spin (1/(larg.mi-haut.)) " << ax[a] << ":  " << xi[1 + a] << '\n';
                    f << "xi_sp/xi_th:  " << xi[1 + a] * M_PI * tem / vF << '\n';
                    f << "xi_sp*T:  " << xi[1 + a] * tem << '\n';
                }
                for (int a = 0; a < 2; a++) {
                    f << "longueur de correlation de spin (formule asymptotique) " << ax[a] << ":  " << xi[3 + a] << '\n';
                    f << "xi_sp_As/xi_th:  " << xi[3 + a] * M_PI * tem / vF << '\n';
                    f << "xi_sp_As*T:  " << xi[3 + a] * tem << '\n';
                }
            }
            if (stages & TPSC_STAGE_TRACE_SG) {
                f << "U*dblOcc-U*n^2/4:   " << U * dblOcc - U * density * density / 4.0 << '\n';
                f << "trace Self*G:  " << w.scalar(TPSC_S_TRSG) << '\n';
            }
            if (stages & TPSC_STAGE_TRACE_SG2) {
                const double sum = w.scalar(TPSC_S_TRSG2);
                f << "trace Self*G2:  " << sum << '\n';
                f << "double occupation (Self*G2): " << (sum + U * density * density / 4.0) / U << '\n';
                f << "mu0:  " << mu0 << '\n';
                f << "mu:  " << mu << '\n';
            }
            f.close();
        }
        if (w.open(f, w.path("./TPSCparams_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat", U, tp, tpp, t3, density, tem,
                             nk0, nbw))) {
            f << std::setprecision(16);
            f << "Usp  " << Usp << "   difRelsp:  " << difRelsp << '\n';
            f << "Uch  " << Uch << "   difRelch:  " << difRelch << '\n';
            f << "nbq  " << nk << '\n';
            f << "nbw  " << nwb << '\n';
            f << "dblOcc_ansatz  " << dblOcc << '\n';
            f << "qxmax  " << q0 << '\n';
            f << "qymax  " << q1 << '\n';
            f << "chispmax_over_chi0max  " << chispmax / chi0max << '\n';
            f << "mu0  " << mu0 << '\n';
            if (stages & TPSC_STAGE_MU) f << std::setprecision(14) << "mu  " << mu << '\n';
            f.close();
        }
        if ((stages & TPSC_STAGE_MU) && w.open(f, w.path("./mu0_mu_D_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat", U,
                                                         tp, tpp, t3, density, tem, nk0, nbw))) {
            f << std::setprecision(15) << std::setiosflags(std::ios::left);
            f << std::setw(25) << mu0 << std::setw(25) << mu << dblOcc << '\n';
        }
    }
    // green::Gwn_local (green.cpp:209-266), which the reference's main runs after find_mu
    if ((stages & TPSC_STAGE_MU) && (stages & TPSC_STAGE_SIGMA)) {
        std::vector<double> gl((size_t)nbw);
        std::ofstream f;
        if (tpsc_gwn_local(c, gl.data()) == TPSC_OK &&
            w.open(f, w.path("./Gwn_loc_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat", U, tp, tpp, t3, density, tem,
                             nk0, nbw))) {
            f << std::setprecision(10) << std::setiosflags(std::ios::left);
            for (int j = 0; j < nbw / 2; j++)
                f << std::setw(10) << j << std::setw(25) << gl[2 * j] << gl[2 * j + 1] << std::endl;
        }
    }
    // what the reference's main runs after find_mu: find_FS_inter (green.cpp:2210-2341) and
    // MDC(n = 0, 1, ..., NP = 10, w = 0, eta = {1e-2, 1e-3, 1e-4, 1e-5}, Pade) (tpsc_main.cpp:77-101,
    // green.cpp:942-1065)
    if ((stages & TPSC_STAGE_MU) && (stages & TPSC_STAGE_SIGMA)) {
        std::vector<int> mp(nk), mq(nk);
        std::vector<double> ap(nk), aq(nk);
        if (tpsc_fs_lattice(c, 1, mp.data(), ap.data()) == TPSC_OK &&
            tpsc_fs_lattice(c, 2, mq.data(), aq.data()) == TPSC_OK) {
            const char* forms[4] = {"./FS_inter_AkF_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                    "./FS_inter_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                    "./FS_inter_AkF_Pade_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                    "./FS_inter_Pade_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat"};
            for (int k = 0; k < 4; k++) {
                std::ofstream f;
                if (!w.open(f, w.path(forms[k], U, tp, tpp, t3, density, tem, nk0, nbw))) continue;
                f << std::setiosflags(std::ios::left) << std::setprecision(8);
                const std::vector<int>& m = (k < 2) ? mp : mq;
                const std::vector<double>& a = (k < 2) ? ap : aq;
                for (int l = 0; l < nk; l++) {
                    f << std::setw(10) << l + m[l] << std::setw(10) << m[l];
                    if (k % 2 == 0) f << a[l];
                    f << std::endl;
                }
            }
        }
        const int NP = std::min(10, nbw / 2), neta = 4;
        const double eta_p[4] = {0.01, 0.001, 0.0001, 0.00001}, wr = 0.0;
        std::vector<int> nq(NP);
        for (int j = 0; j < NP; j++) nq[j] = j;
        const long nk8 = (long)nk * (nk + 1) / 2;
        std::vector<double> sw(2 * (size_t)nk8 * neta), akw((size_t)nk8 * neta);
        if (NP >= 2 && tpsc_mdc(c, nq.data(), NP, wr, eta_p, neta, 1, sw.data(), akw.data()) == TPSC_OK) {
            std::ofstream fs, fa;
            if (w.open(fs, w.path("Self_Pade_k_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%5.3e_%1d_%1d.dat", U, tp, tpp, t3,
                                  density, tem, wr, nk0, nbw)) &&
                w.open(fa, w.path("Akw_Pade_k_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%5.3e_%1d_%1d.dat", U, tp, tpp, t3,
                                  density, tem, wr, nk0, nbw))) {
                fs << std::setprecision(12) << std::setiosflags(std::ios::left);
                fa << std::setprecision(12) << std::setiosflags(std::ios::left);
                for (int l = 0; l < nk; l++)
                    for (int m = 0; m <= l; m++) {
                        const long k = m + (long)l * (l + 1) / 2;
                        fs << std::setw(10) << l << std::setw(20) << m;
                        fa << std::setw(10) << l << std::setw(20) << m;
                        for (int e = 0; e < neta; e++) {
                            fs << std::setw(20) << sw[2 * (k * neta + e)] << std::setw(25) << sw[2 * (k * neta + e) + 1];
                            fa << std::setw(20) << akw[k * neta + e];
                        }
                        fs << std::endl;
                        fa << std::endl;
                    }
            }
        }
    }
    // cond_opt::calc_chijj_vertex_corr_all_optim (cond_opt.cpp:4195-4279)
    if (stages & TPSC_STAGE_CHIJJ) {
        const double kx0 = w.scalar(TPSC_S_KX0);
        const std::vector<double> j1 = w.field(TPSC_F_CHIJJ1), j2 = w.field(TPSC_F_CHIJJ2);
        std::ofstream f;
        if (w.open(f, w.path("kx0_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat", U, tp, tpp, t3, density, tem, nk0, nbw))) {
            f << std::setprecision(16) << std::setiosflags(std::ios::left) << std::setiosflags(std::ios::scientific) << kx0
              << '\n';
            f.close();
        }
        if (w.open(f, w.path("kx0_bin_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat", U, tp, tpp, t3, density, tem, nk0,
                             nbw),
                   std::ios::out | std::ios::binary)) {
            f.write((const char*)&kx0, sizeof(kx0));
            f.close();
        }
        const char* forms[3] = {"./chijj1_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                "./chijj_with_vc1_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                "./chijj2_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat"};
        for (int k = 0; k < 3 && !j1.empty() && j1.size() == j2.size(); k++) {
            if (!w.open(f, w.path(forms[k], U, tp, tpp, t3, density, tem, nk0, nbw))) continue;
            f << std::setprecision(16) << std::setiosflags(std::ios::left) << std::setiosflags(std::ios::scientific);
            for (int j = 0; j <= nbw / 2; j++) {
                const double v = (k == 0) ? j1[j] : (k == 2) ? j2[j] : j1[j] + j2[j];
                f << std::setw(10) << j << v << '\n';
            }
            f.close();
        }
        // the angle-binned contributions and their sums (cond_opt.cpp:3889-3946)
        const int nth = (int)w.scalar(TPSC_S_N_THETA);
        const std::vector<double> k1 = w.field(TPSC_F_CHIJJ1_K), k2 = w.field(TPSC_F_CHIJJ2_K);
        const char* kforms[4] = {"./chijj1_k_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                 "./chijj1_k_tot_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                 "./chijj2_k_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat",
                                 "./chijj2_k_tot_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat"};
        for (int k = 0; k < 4 && nth >= 1 && k1.size() == (size_t)nth * (nbw / 2 + 1); k++) {
            if (!w.open(f, w.path(kforms[k], U, tp, tpp, t3, density, tem, nk0, nbw))) continue;
            f << std::setprecision(16) << std::setiosflags(std::ios::left) << std::setiosflags(std::ios::scientific);
            const std::vector<double>& b = (k < 2) ? k1 : k2;
            const std::vector<double>& qn = (k < 2) ? j1 : j2;
            for (int j = 0; j <= nbw / 2; j++) {
                if (k % 2 == 1) {
                    double tot = 0.0;
                    for (int p = 0; p < nth; p++) tot += b[p + (size_t)j * nth];
                    f << std::setw(10) << j << std::setw(30) << tot << std::setw(30) << qn[j] << tot / qn[j] << '\n';
                } else
                    for (int p = 0; p < nth; p++) f << std::setw(10) << j << std::setw(10) << p + 1 << b[p + (size_t)j * nth] << '\n';
            }
            f.close();
        }
    }
    // Aslamazov-Larkin outputs (cond_opt.cpp:6447-6481)
    if (stages & TPSC_STAGE_AL) {
        const int mmin = (int)w.scalar(TPSC_S_AL_MMIN);
        const std::vector<double> fr = w.field(TPSC_F_AL_FREQ), j3 = w.field(TPSC_F_CHIJJ3),
                                  ja = w.field(TPSC_F_CHIJJ_ALL);
        const char* forms[2] = {"./chijj3_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d_%1d.dat",
                                "./chijj_with_all_vc_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d_%1d.dat"};
        for (int k = 0; k < 2 && !fr.empty(); k++) {
            std::ofstream f;
            if (!w.open(f, w.path(forms[k], U, tp, tpp, t3, density, tem, nk0, nbw, mmin))) continue;
            f << std::setprecision(16) << std::setiosflags(std::ios::left) << std::setiosflags(std::ios::scientific);
            for (size_t j = 0; j < fr.size(); j++) f << std::setw(10) << (long)fr[j] << (k == 0 ? j3[j] : ja[j]) << '\n';
        }
        // with N_theta_k > 1, by angle bin (cond_opt.cpp:6483-6520)
        const int nth = (int)w.scalar(TPSC_S_N_THETA);
        const std::vector<double> k3 = (nth > 1) ? w.field(TPSC_F_CHIJJ3_K) : std::vector<double>();
        const char* kforms[2] = {"./chijj3_k_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d_%1d.dat",
                                 "./chijj3_k_tot_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d_%1d.dat"};
        for (int k = 0; k < 2 && nth > 1 && k3.size() == (size_t)nth * fr.size(); k++) {
            std::ofstream f;
            if (!w.open(f, w.path(kforms[k], U, tp, tpp, t3, density, tem, nk0, nbw, mmin))) continue;
            f << std::setprecision(16) << std::setiosflags(std::ios::left) << std::setiosflags(std::ios::scientific);
            for (size_t j = 1; j < fr.size(); j++) {
                if (k == 1) {
                    double tot = 0.0;
                    for (int p = 0; p < nth; p++) tot += k3[p + j * nth];
                    f << std::setw(10) << (long)fr[j] << std::setw(30) << tot << std::setw(30) << j3[j] << tot / j3[j] << '\n';
                } else
                    for (int p = 0; p < nth; p++) f << std::setw(10) << (long)fr[j] << std::setw(10) << (p + 1) << k3[p + j * nth] << '\n';
            }
        }
    }
    if (nfiles) *nfiles = w.count;
    return w.err;
}
