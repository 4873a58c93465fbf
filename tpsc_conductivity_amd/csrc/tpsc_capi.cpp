// tpsc_capi.cpp -- host side of libtpsc_hip.so: the C ABI of include/tpsc_capi.h over the
// HIP kernels of tpsc_kernels.hip.  Stage order and scalar logic follow the reference
// object methods (cond_opt > green > chi > chi0 > green0 > hamiltonien); the heavy
// arrays stay in HBM from chi0 to Lambda_jj, only scalars cross to the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../include/tpsc_capi.h"
#include "tpsc_launch.hpp"
#include "tpsc_queue.hpp"

using namespace tpsc;

const char* tpsc_noalloc_error();  // tpsc_noalloc.hip: message of the last failed call on this thread

namespace {

struct ProfPair {
    std::string name;
    hipEvent_t a, b;
    hipStream_t s;
    bool open;  // b not yet recorded
};

// host_scalar: [0..2] reductions read by wait_mapped, [4] Sigma_inf, [8..15] root-search mirror,
// [16..18] mu0 search {found, mu0, evaluations}, [19] chi0max, [20] its wedge index (int),
// [21..22] Tr(Sigma G) main sum and correction, [23] Tr(Sigma G) at the final mu
constexpr int kHostScalars = 32;
constexpr int kMaxTheta = 64;  // angle bins of tpsc_chijj (N_theta_k)
// Mixed path: chi_sp = chi0/(1 - Usp chi0/2) carries the relative error of chi0 multiplied by the
// spin enhancement 1/(1 - Usp chi0max/2).  FP32 tau arrays leave ~6e-8 in chi0, so above this
// enhancement the chi0 stage runs in FP64 (measured: 3.9e-7 on chi_sp at 7.2, 1.1e-6 at ~12).
constexpr double kPromoteEnhancement = 8.0;

}  // namespace

struct tpsc_ctx {
    tpsc_params p{};
    // current grid (green0::set_grid)
    int nbk = 0, nbw = 0, nwmax = 0;
    Grid g{};
    double mu0 = 0.0;
    double nsr = 0.0;
    // results
    double chi0max = 0.0;
    int qmax[2] = {0, 0};
    double Usp = 0.0, Uch = 0.0, dblOcc = 0.0, difRelsp = 0.0, difRelch = 0.0;
    double Sigma_inf = 0.0, trSG = 0.0, trSG_corr = 0.0, trSG2 = 0.0, dblOcc2 = 0.0;
    double mu = 0.0;
    bool mu_ok = false;
    int nev_sp = 0, nev_ch = 0, nev_mu = 0;
    int roots_dev = 0;  // searches of the last solve settled on the device: 1 Usp, 2 Uch, 4 mu
    std::vector<double> chijj1, chijj2;
    int ntheta = 1;                            // angle bins of the last Lambda_jj call
    std::vector<double> chijj1_k, chijj2_k;    // angle-binned Lambda_jj1, Lambda_jj2, [b + n*ntheta]
    std::vector<double> chijj3_k;              // angle-binned Lambda_jj3 on the sparse set
    int* jjbins = nullptr;                     // angle bin of each Lambda point (ntheta > 1)
    int jjbins_ntheta = 0;
    double* jjpart_t = nullptr;                // binned partial rows and their reduction (ntheta > 1)
    double* jjred_t = nullptr;
    size_t jjpart_t_cap = 0, jjred_t_cap = 0;
    double* al_red_t = nullptr;                // binned Lambda_jj3 (ntheta > 1)
    size_t al_red_t_cap = 0;
    std::vector<long> al_freq;                 // Aslamazov-Larkin sparse boson set
    std::vector<double> chijj3, chijj_all;     // Lambda_jj3 and Lambda_jj1+2+3 on it
    double kx0 = 0.0, jj_inf1 = 0.0, jj_inf2 = 0.0, jj_inf4 = 0.0;
    // stage state
    bool have_chi0 = false, have_U = false, have_V = false, have_sigma = false, have_jj = false;
    bool have_trSG = false, have_trSG2 = false;
    int al_mmin = -1;
    double* gloc_buf = nullptr;  // k_gloc partials + result, allocated at the first tpsc_gwn_local
    bool mu0_fixed = false;      // tpsc_fix_mu0: mu0 is an input, not solved from n
    bool mu0_failed = false;     // the last mu0 search found no root (mu0 kept its previous value)
    TanArrays tan{};             // tangent arrays of the Sigma-loss gradient
    bool grad_ready = false;     // forward of the loss done, weights on the device
    int loss_terms = 0;          // TPSC_LOSS_* of the last tpsc_loss_forward
    std::vector<double> grad_terms, grad_dmu, grad_dlambda;  // per-parameter tangents of the last backward
    bool keep_chi_tau = false;
    bool chi_tau_fp64 = false;   // the kept chi0(q,tau) came from an FP64 chi0 stage
    // mixed path: the chi0 stage of the full grid runs in FP64 (chi0(q,tau) in chiw, chi0(r,tau) in
    // chirt64, then narrowed into chirt) once the spin enhancement of the point is found large
    bool chi0_fp64 = false;
    double* chirt64 = nullptr;
    // device
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // boundary kernels, overlapped with the slab kernels
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_tau0 = nullptr, ev_hf0 = nullptr;
    bool hf0_pending = false;  // h(r,tau), f0(r,tau) queued on the prep stream at the chi0 stage
    // third stream: the Lambda stage's hV/f0V slabs, transposes and boundary data, queued at the
    // end of the vertex stage and overlapped with the self-energy and mu stages
    hipStream_t prep = nullptr;
    // captured chunks of root-search passes, one per (grid, precision, search kind), made once
    // the chunk has been launched TPSC_GRAPH_AFTER times (default 16)
    struct RootGraph {
        int N, Nw, prec, kind, uses;
        hipGraphExec_t exec;
    };
    std::vector<RootGraph> root_graphs;
    hipEvent_t ev_hv = nullptr, ev_v = nullptr;
    hipEvent_t ev_bnd = nullptr;  // the Lambda boundary kernel (hVb, f0Vb, slopes) done on prep
    bool hv_pending = false;
    bool hv_slabs_due = false;  // the hV, f0V slabs are still to be queued (after the Sigma slab)
    bool hv_dropped = false;    // they were dropped: the Lambda stage makes them
    hipEvent_t ev_sig = nullptr;
    bool sinf_async = false;  // Sigma_inf copied to host_scalar[4], not yet read
    // mu0 is found on the device and read there by the chi0 and h/f0 kernels; the host copy
    // (host_scalar[16..18]) is taken when it is first needed, after ev_mu0
    hipEvent_t ev_mu0 = nullptr;
    bool mu0_async = false;
    // chi0max and its index, copied to host_scalar[19..20] behind the chi0 stage, taken after
    // the first round trip of the Usp search, whose bracket the device sets from chi0max
    bool chi0_async = false;
    int trace_async = 0;  // trace reductions queued, not yet read: 1 Tr(Sigma G), 2 at the final mu
    DevArrays d{};
    ALArrays al{};
    double* chi_tau = nullptr;
    double* host_scalar = nullptr;  // pinned, mapped (kHostScalars): kernels write root-finder values here
    double* dev_scalar = nullptr;   // device alias of host_scalar
    size_t dev_bytes = 0;
    std::vector<void*> allocs;
    std::string err;
    // profiling
    bool prof = false;
    std::vector<ProfPair> pending;
    std::map<std::string, std::pair<double, int>> prof_acc;
    std::vector<std::string> prof_names;
};

namespace {

// nearest-neighbour hopping: tpsc_params.t, with 0 meaning the reference's fixed value 1
double hop_t(const tpsc_ctx* c) { return c->p.t != 0.0 ? c->p.t : 1.0; }

int fail(tpsc_ctx* c, int code, const std::string& msg)
{
    c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return fail((c), TPSC_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

void prof_begin(tpsc_ctx* c, const char* name, hipStream_t st = nullptr)
{
    if (!c->prof) return;
    ProfPair pp;
    pp.name = name;
    pp.s = st ? st : c->stream;
    (void)hipEventCreate(&pp.a);
    (void)hipEventCreate(&pp.b);
    (void)hipEventRecord(pp.a, pp.s);
    pp.open = true;
    c->pending.push_back(pp);
}
void prof_end(tpsc_ctx* c)
{
    if (!c->prof || c->pending.empty() || !c->pending.back().open) return;
    (void)hipEventRecord(c->pending.back().b, c->pending.back().s);
    c->pending.back().open = false;
}
// the closed regions (all three streams drained first); a region still open stays pending
void prof_flush(tpsc_ctx* c)
{
    if (c->pending.empty()) return;
    (void)hipStreamSynchronize(c->prep);
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamSynchronize(c->stream);
    std::vector<ProfPair> keep;
    for (auto& pp : c->pending) {
        if (pp.open) {
            keep.push_back(pp);
            continue;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, pp.a, pp.b);
        auto it = c->prof_acc.find(pp.name);
        if (it == c->prof_acc.end()) {
            c->prof_acc[pp.name] = {ms, 1};
            c->prof_names.push_back(pp.name);
        } else {
            it->second.first += ms;
            it->second.second += 1;
        }
        (void)hipEventDestroy(pp.a);
        (void)hipEventDestroy(pp.b);
    }
    c->pending.swap(keep);
}

// ------------------------------------------------------------------ root finder
// the derivative-free bracketed search of tpsc_roots.hpp over a host-evaluated function, in the
// reference's calling form {start, lo, hi} + optional limits
bool find_zero(const std::function<double(double)>& f, const double init[3], double* root, const double* lims,
               int* nevals)
{
    return find_root(f, init, root, lims, nevals);
}

// TPSC_TRACE_ROOTS=1 prints every root-finder pass to stderr
bool trace_roots()
{
    static const bool on = getenv("TPSC_TRACE_ROOTS") && atoi(getenv("TPSC_TRACE_ROOTS")) != 0;
    return on;
}

// ------------------------------------------------------------------ device helpers
// count elements of elem bytes (elem = 0: sizeof(T))
template <typename T>
int dalloc(tpsc_ctx* c, T** ptr, size_t count, size_t elem = 0)
{
    void* p = nullptr;
    const size_t bytes = (elem ? elem : sizeof(T)) * (count ? count : 1);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, TPSC_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    c->allocs.push_back(p);
    c->dev_bytes += bytes;
    *ptr = (T*)p;
    return 0;
}

// bytes of one element of the tau / iqn arrays (FP64 path: 8, mixed path: 4)
size_t elem_bytes(const tpsc_ctx* c) { return c->p.precision ? sizeof(float) : sizeof(double); }

int read_scalar(tpsc_ctx* c, const double* dptr, double* out)
{
    HIPCHK(c, hipMemcpyAsync(c->host_scalar + 1, dptr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->host_scalar[1];
    return 0;
}

// a reduction launched with out = dev_scalar has written its value(s) to mapped host memory
int wait_mapped(tpsc_ctx* c, double* out, int n = 1)
{
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; i++) out[i] = ((volatile double*)c->host_scalar)[i];
    return 0;
}

// Halley root search (tpsc_roots.hpp) run on the device: pass() queues one evaluation and the
// step kernel that advances the state in d.root; kRootChunk passes are queued per host round
// trip (a settled search turns the remaining ones into early exits).  The outcome is read from
// the mirror in mapped host memory, host_scalar[8..15], which says "running" until it settles.
// A full chunk of a grid, precision and search kind that has already been launched
// TPSC_GRAPH_AFTER times (default 16) is captured into a graph and replayed from
// then on: the pass and step kernels see the point only through d.root.  Capture and
// instantiation cost ~1.2 ms per graph and a replay saves ~15 us per chunk on small grids, so
// only a grid that keeps being solved gets one.  TPSC_NO_GRAPHS=1 never captures.
// Five: most searches settle in four or five passes, and each pass queued after that costs ~8 us.
constexpr int kRootChunk = 5;
enum { kRootUsp = 0, kRootUch = 1, kRootMu = 2 };
int root_chunk_graph(tpsc_ctx* c, int kind, const std::function<hipError_t()>& pass, hipGraphExec_t* out)
{
    static const int after = getenv("TPSC_GRAPH_AFTER") ? atoi(getenv("TPSC_GRAPH_AFTER")) : 16;
    const Grid& g = c->g;
    *out = nullptr;
    tpsc_ctx::RootGraph* rg = nullptr;
    for (auto& r : c->root_graphs)
        if (r.N == g.N && r.Nw == g.Nw && r.prec == g.prec && r.kind == kind) rg = &r;
    if (!rg) {
        c->root_graphs.push_back({g.N, g.Nw, g.prec, kind, 0, nullptr});
        rg = &c->root_graphs.back();
    }
    if (rg->exec) {
        *out = rg->exec;
        return 0;
    }
    if (rg->uses++ < after) return 0;  // launched one by one until then
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < kRootChunk && e == hipSuccess; k++) e = pass();
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (e == hipSuccess) e = ec;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(c, TPSC_E_DEVICE, std::string("root-search graph: ") + hipGetErrorString(e));
    rg->exec = exec;
    *out = exec;
    return 0;
}

int mu0_ready(tpsc_ctx* c);
// chi0max (device) != nullptr: the bracket is (0, 2/chi0max - 2 eps), set on the device.
int device_halley(tpsc_ctx* c, int kind, const std::function<hipError_t()>& pass, HalleyState s0, double* root,
                  int* nevals, bool* ok, const double* chi0max = nullptr)
{
    static const bool graphs = !(getenv("TPSC_NO_GRAPHS") && atoi(getenv("TPSC_NO_GRAPHS")) != 0);
    if (kind == kRootMu)
        if (int rc = mu0_ready(c)) return rc;
    s0.T = c->g.T;
    s0.mu0 = c->mu0;
    HIPCHK(c, launch_halley_init(c->d.root, s0, chi0max, c->dev_scalar + 8, c->stream));
    volatile const double* m = c->host_scalar + 8;
    int issued = 0;
    for (;;) {
        const int n = std::min(kRootChunk, s0.itmax - issued);
        hipGraphExec_t exec = nullptr;
        if (graphs && n == kRootChunk) {
            int rc = root_chunk_graph(c, kind, pass, &exec);
            if (rc) return rc;
        }
        if (exec)
            HIPCHK(c, hipGraphLaunch(exec, c->stream));
        else
            for (int k = 0; k < n; k++) HIPCHK(c, pass());
        issued += n;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (trace_roots() && (int)m[0] != kRootRunning)
            fprintf(stderr, "[tpsc root] settled after %d passes: x=%.17g f=%.6e f1=%.6e [%.17g, %.17g] status %d\n",
                    (int)m[3], m[2], m[4], m[5], m[6], m[7], (int)m[0]);
        if ((int)m[0] != kRootRunning || issued >= s0.itmax) break;
    }
    *ok = (int)m[0] == kRootConverged;
    *root = m[1];
    if (nevals) *nevals = (int)m[3];
    return 0;
}

// the side stream waits for the Lambda-stage work queued on the prep stream (which reads and
// rewrites h, f0 and the transposed rows); nothing is left pending afterwards
int queue_hv_slabs(tpsc_ctx* c);
// need_slabs: the caller is the Lambda stage, which wants the hV, f0V slabs (queued now if they
// were not yet); any other caller drops slabs that were not queued, and the Lambda stage then
// makes them itself
int settle_prep(tpsc_ctx* c, bool need_slabs)
{
    if (!c->hv_pending) return 0;
    if (need_slabs) {
        if (int rc = queue_hv_slabs(c)) return rc;
    } else if (c->hv_slabs_due) {
        c->hv_slabs_due = false;
        c->hv_dropped = true;
    }
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_hv, 0));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_hv, 0));
    c->hv_pending = false;
    return 0;
}

// the host value of Sigma_inf (its copy was queued by the self-energy stage)
int sigma_inf_ready(tpsc_ctx* c)
{
    if (!c->sinf_async) return 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->Sigma_inf = ((volatile double*)c->host_scalar)[4];
    c->sinf_async = false;
    return 0;
}

// the side stream starts after the work queued so far on the main stream
int fork_side(tpsc_ctx* c)
{
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    return 0;
}
// the main stream continues after the work queued so far on the side stream
int join_side(tpsc_ctx* c)
{
    HIPCHK(c, hipEventRecord(c->ev_join, c->side));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return 0;
}

// green0::find_mu (green0.cpp:240-262) on the device: the whole search in one launch
// (Halley steps on the device density sums), queued without a wait.  The root lands in d.mu0v
// for the kernels and in host_scalar[16..18] for mu0_ready.  A fixed mu0, or mu0 = 0 at half
// filling without t', t'' (green0.cpp:245), is stored as a preset.
int find_mu0(tpsc_ctx* c)
{
    double preset = std::nan("");
    if (c->mu0_fixed)
        preset = c->mu0;
    else if (c->p.n == 1.0 && c->p.tp == 0 && c->p.tpp == 0)
        preset = 0.0;
    HIPCHK(c, launch_mu0_root(c->g, c->d, c->p.n, preset, c->dev_scalar + 16, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_mu0, c->stream));
    c->mu0_async = true;
    return 0;
}

// the host copy of mu0 once the search has ended
int mu0_ready(tpsc_ctx* c)
{
    if (!c->mu0_async) return 0;
    HIPCHK(c, hipEventSynchronize(c->ev_mu0));
    volatile const double* v = c->host_scalar + 16;
    c->mu0_async = false;
    c->mu0_failed = v[0] == 0.0;
    if (c->mu0_failed)  // the previous mu0 is kept, as in the reference ("mu0 non trouve")
        return fail(c, TPSC_E_NOROOT, "mu0 not found: the density equation has no root; mu0 keeps its previous value");
    c->mu0 = v[1];
    return 0;
}

// Drops every computed stage and result (chi0 through Lambda_jj, traces, gradients), after the
// Lambda-stage work still queued on the prep stream (it reads eps(k), mu0, h and f0) has ended.
// The device arrays are kept.
int invalidate_stages(tpsc_ctx* c)
{
    if (c->hv_pending || c->hf0_pending) {
        HIPCHK(c, hipStreamSynchronize(c->prep));
        c->hf0_pending = false;
        c->hv_pending = false;
    }
    c->have_chi0 = c->have_U = c->have_V = c->have_sigma = c->have_jj = false;
    c->have_trSG = c->have_trSG2 = false;
    c->trace_async = 0;
    c->mu_ok = false;
    c->grad_ready = false;
    c->hv_slabs_due = c->hv_dropped = false;
    c->chijj1.clear();
    c->chijj2.clear();
    c->chijj1_k.clear();
    c->chijj2_k.clear();
    c->chijj3_k.clear();
    c->chijj3.clear();
    c->chijj_all.clear();
    c->al_freq.clear();
    c->al_mmin = -1;
    return 0;
}

// green0::set_grid (green0.h:44): quadrant eps(k), wedge tables, mu0
int set_grid(tpsc_ctx* c, int N, int Nw)
{
    if (!supported_N(N) || !supported_Nw(Nw))
        return fail(c, TPSC_E_INVALID, "unsupported grid N=" + std::to_string(N) + " Nw=" + std::to_string(Nw));
    const int nk = N / 2 + 1;
    c->nbk = nk;
    c->nbw = Nw;
    c->nwmax = Nw / 2;
    Grid& g = c->g;
    g.N = N;
    g.nk = nk;
    g.Nw = Nw;
    g.nwb = Nw / 2 + 1;
    g.nk8 = (long)nk * (nk + 1) / 2;
    g.nq = (long)nk * (nk - 2);
    g.T = c->p.T;
    g.prec = c->p.precision;
    hipStream_t s = c->stream;
    HIPCHK(c, launch_model(g, c->d, hop_t(c), c->p.tp, c->p.tpp, c->p.t3, s));
    HIPCHK(c, launch_fill_tw(c->d, Nw, s));
    c->have_chi0 = false;
    return find_mu0(c);
}

// ------------------------------------------------------------------ stages
bool full_grid(const tpsc_ctx* c) { return c->g.N == c->p.nk0 && c->g.Nw == c->p.Nw0; }

// chi0::calc_chiqw_FFT_spline, chi0.cpp:279-718
int chi0_ready(tpsc_ctx* c);
// defer: chi0max and its index are copied to the host behind the stage without a wait, and
// chi0_ready takes them (the Usp search, whose bracket the device sets, syncs first)
int stage_chi0(tpsc_ctx* c, int nq0, int nw0, bool defer = false)
{
    if (nq0 > c->p.nk0 || nw0 > c->p.Nw0) return fail(c, TPSC_E_INVALID, "chi0 grid larger than the context grid");
    // Lambda-stage work of an earlier point (prep stream) still reads eps(k), V, h and f0
    int rc = settle_prep(c, false);
    if (rc) return rc;
    hipStream_t s = c->stream;
    if (c->hf0_pending) {  // h, f0 of an earlier chain, still read on the prep stream, use eps(k)
        HIPCHK(c, hipStreamWaitEvent(s, c->ev_hf0, 0));
        c->hf0_pending = false;
    }
    rc = set_grid(c, nq0, nw0);
    if (rc) return rc;
    // the arrays and precision this stage computes in: FP64 throughout for a promoted full grid
    // of the mixed path (chi0(q,tau) transformed in place in chiw, chi0(r,tau) in chirt64)
    const bool f64 = c->p.precision && c->chi0_fp64 && full_grid(c);
    Grid gx = c->g;
    DevArrays dx = c->d;
    if (f64) {
        if (!c->chirt64 && (rc = dalloc(c, &c->chirt64, (size_t)c->g.nwb * c->g.nk8))) return rc;
        gx.prec = 0;
        dx.chi = c->d.chiw;
        dx.chirt = c->chirt64;
    }
    // Nothing waits for mu0 on the host.  The tau = 0 slice (one workgroup of 1024 threads)
    // follows the mu0 search in the main stream, so that it takes its CU before the slab kernel
    // on the side stream fills the chip: alone it runs ~240 us at N = 256, behind the slab
    // kernel's workgroups more than twice that, and the chi0 pencils wait for it.
    if ((rc = fork_side(c))) return rc;
    prof_begin(c, "chi0_slab", c->side);
    HIPCHK(c, launch_chi0_slab(gx, dx, c->side));
    prof_end(c);
    prof_begin(c, "chi0_tau0");
    HIPCHK(c, launch_chi0_tau0(gx, dx, s));
    prof_end(c);
    HIPCHK(c, hipEventRecord(c->ev_tau0, s));
    if ((rc = join_side(c))) return rc;
    // h(r,tau) and f0(r,tau) of the Lambda stage need only mu0 and eps(k): on the full grid
    // they are queued now on the prep stream, behind the tau = 0 slice, and run beside the chi0
    // slab and pencils and the sum-rule passes; the Lambda stage waits for ev_hf0
    if (full_grid(c)) {
        HIPCHK(c, hipStreamWaitEvent(c->prep, c->ev_tau0, 0));
        prof_begin(c, "jj_hf0_slab", c->prep);
        HIPCHK(c, launch_hf0_slab(c->g, c->d, c->prep));
        prof_end(c);
        HIPCHK(c, hipEventRecord(c->ev_hf0, c->prep));
        c->hf0_pending = true;
    }
    if (c->keep_chi_tau && c->chi_tau) {
        HIPCHK(c, hipMemcpyAsync(c->chi_tau, dx.chi, (f64 ? sizeof(double) : elem_bytes(c)) * c->g.nwb * c->g.nk8,
                                 hipMemcpyDeviceToDevice, s));
        c->chi_tau_fp64 = f64;
    }
    prof_begin(c, "chi0_pencil");
    HIPCHK(c, launch_chi0_pencil(gx, dx, s));
    prof_end(c);
    if (f64) HIPCHK(c, launch_narrow(c->chirt64, (float*)c->d.chirt, (long)c->g.nwb * c->g.nk8, s));
    prof_begin(c, "sr_moments");
    HIPCHK(c, launch_sr_moments(c->g, c->d, s));
    prof_end(c);
    HIPCHK(c, launch_chi0max(c->g, c->d, s));
    HIPCHK(c, hipMemcpyAsync(c->host_scalar + 19, c->d.red, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(c->host_scalar + 20, c->d.imax, sizeof(int), hipMemcpyDeviceToHost, s));
    c->chi0_async = true;
    if (defer) return 0;
    HIPCHK(c, hipStreamSynchronize(s));
    return chi0_ready(c);
}

int chi0_ready(tpsc_ctx* c)
{
    if (!c->chi0_async) return 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_flush(c);
    volatile const double* h = c->host_scalar;
    c->chi0max = h[19];
    const int imax = *(volatile const int*)(h + 20);
    int l = 0;
    while ((long)(l + 1) * (l + 2) / 2 <= imax) l++;
    c->qmax[0] = l;
    c->qmax[1] = imax - l * (l + 1) / 2;
    c->chi0_async = false;
    if (int rc = mu0_ready(c)) return rc;
    if (c->mu0_failed) return fail(c, TPSC_E_NOROOT, "chi0 computed at a mu0 whose search failed");
    c->have_chi0 = true;
    return 0;
}

int eval_sum_rule(tpsc_ctx* c, int channel, double Ux, double* out)
{
    if (!c->have_chi0) return fail(c, TPSC_E_ORDER, "sum rule before chi0");
    double tr = 0.0;
    HIPCHK(c, launch_sum_rule(c->g, c->d, channel == 0 ? -Ux : Ux, c->dev_scalar, c->stream));
    int rc = wait_mapped(c, &tr);
    if (rc) return rc;
    const double n = c->p.n, nsr = c->nsr, U = c->p.U;
    if (channel == 0)
        *out = tr + n - n * n / 2.0 - nsr + Ux * nsr * nsr / (2.0 * U);
    else
        *out = tr + n - n * n / 2.0 - (nsr + c->Usp * nsr * nsr / (2.0 * U) - nsr * nsr);
    return 0;
}

// chi::find_Usp_Uch(nq0, nw0, Usc), chi.cpp:846-876
int stage_UspUch_grid(tpsc_ctx* c, int nq0, int nw0, double Usc[2], bool* ok_out)
{
    int rc = stage_chi0(c, nq0, nw0, true);
    if (rc) return rc;
    prof_begin(c, "sum_rules");
    const double U = c->p.U;
    auto fsp = [&](double x) {
        double v = 0.0;
        if (rc == 0) rc = eval_sum_rule(c, 0, x, &v);
        return v;
    };
    auto fch = [&](double x) {
        double v = 0.0;
        if (rc == 0) rc = eval_sum_rule(c, 1, x, &v);
        return v;
    };
    // sum rule, d/dUx and d2/dUx2 in one device pass; channel 0: sU = -Ux, channel 1: sU = +Ux.
    // The host-known part of each channel's function is a0 + a1 Ux.
    const double n = c->p.n, nsr = c->nsr;
    auto pass_sp = [&]() { return launch_sum_rule3_pass(c->g, c->d, -1.0, c->dev_scalar + 8, c->stream); };
    auto pass_ch = [&]() { return launch_sum_rule3_pass(c->g, c->d, 1.0, c->dev_scalar + 8, c->stream); };
    double root = 0.0;
    bool ok = true;
    int nh = 0;
    bool hok = false;
    // bracket (0, 2/chi0max - 2 eps) with the reference's start U/2.  That start can lie beyond
    // the Stoner bound 2/chi0max, where the function has its pole: the derivative-based search
    // starts at the midpoint instead (same root).  The device sets both from its chi0max, so the
    // first passes follow the chi0 stage without a host round trip.
    rc = device_halley(c, kRootUsp, pass_sp,
                       halley_start(U / 2.0, 0.0, 0.0, 40, n - n * n / 2.0 - nsr, nsr * nsr / (2.0 * U)), &root,
                       &nh, &hok, c->d.red);
    if (!rc) rc = chi0_ready(c);
    if (rc) return rc;
    double init[3] = {U / 2.0, 0.0, 2.0 / c->chi0max - 2 * DBL_EPSILON};
    c->roots_dev = (c->roots_dev & ~1) | (hok ? 1 : 0);
    if (!hok) hok = find_zero(fsp, init, &root, nullptr, &c->nev_sp);
    else c->nev_sp = nh;
    if (rc) return rc;
    if (hok) {
        Usc[0] = root;
        c->Usp = root;
    } else
        ok = false;
    root = 0.0;
    init[0] = 0;
    init[2] = 1000.0 * U;
    rc = device_halley(c, kRootUch, pass_ch,
                       halley_start(init[0], init[1], init[2], 40,
                                    n - n * n / 2.0 - (nsr + c->Usp * nsr * nsr / (2.0 * U) - nsr * nsr), 0.0),
                       &root, &nh, &hok);
    if (rc) return rc;
    c->roots_dev = (c->roots_dev & ~2) | (hok ? 2 : 0);
    if (!hok) hok = find_zero(fch, init, &root, nullptr, &c->nev_ch);
    else c->nev_ch = nh;
    if (rc) return rc;
    if (hok)
        Usc[1] = root;
    else
        ok = false;
    prof_end(c);
    prof_flush(c);
    if (rc) return rc;
    *ok_out = ok;
    return 0;
}

// chi::find_Usp_Uch(), chi.cpp:210-417
int stage_UspUch(tpsc_ctx* c)
{
    double Usc[2] = {0.0, 0.0}, Us0 = 0.0, Uc0 = 0.0;
    bool ok = true, ok0 = true;
    int rc;
    if (c->p.coarse_pass) {
        rc = stage_UspUch_grid(c, c->p.nk0 / 2, c->p.Nw0 / 2, Usc, &ok0);
        if (rc) return rc;
        Us0 = Usc[0];
        Uc0 = Usc[1];
    }
    rc = stage_UspUch_grid(c, c->p.nk0, c->p.Nw0, Usc, &ok);
    if (rc) return rc;
    if (c->p.precision && !c->chi0_fp64 && ok && 1.0 / (1.0 - Usc[0] * c->chi0max / 2.0) > kPromoteEnhancement) {
        c->chi0_fp64 = true;  // kept for this point: later solves start in FP64
        rc = stage_UspUch_grid(c, c->p.nk0, c->p.Nw0, Usc, &ok);
        if (rc) return rc;
    }
    c->difRelsp = fabs((Usc[0] - Us0) / Usc[0]);
    c->difRelch = fabs((Usc[1] - Uc0) / Usc[1]);
    c->Usp = Usc[0];
    c->Uch = Usc[1];
    const double U = c->p.U, nsr = c->nsr;
    c->dblOcc = (c->p.n > 1.0) ? c->Usp / U * nsr * nsr / 4.0 + 1.0 - nsr : c->Usp / U * nsr * nsr / 4.0;
    c->have_U = true;
    if (!ok) return fail(c, TPSC_E_NOROOT, "Usp or Uch root not found");
    return 0;
}


// chi::calc_chirt, chi.cpp:37-206
int stage_vertex(tpsc_ctx* c)
{
    if (!c->have_U || !c->have_chi0 || !full_grid(c)) {
        int rc = stage_UspUch(c);
        if (rc) return rc;
    }
    int rc = settle_prep(c, false);  // queued Lambda work still reads V
    if (rc) return rc;
    prof_begin(c, "vertex_slab");
    HIPCHK(c, launch_vertex_slab(c->g, c->d, c->p.U, c->Usp, c->Uch, c->stream));
    prof_end(c);
    prof_begin(c, "vertex_pencil");
    HIPCHK(c, launch_vertex_pencil(c->g, c->d, c->p.U, c->Usp, c->Uch, c->stream));
    prof_end(c);
    HIPCHK(c, hipEventRecord(c->ev_v, c->stream));  // V and V(r,0) are complete here
    prof_flush(c);
    c->have_V = true;
    // The Lambda stage's boundary data, hV and f0V slabs (h, f0 times V) and their point-major
    // transposes need V, mu0, Usp, Uch and the h, f0 slabs of the chi0 stage only.  They run on
    // the prep stream: the boundary kernel (four workgroups) from here, the slabs once the
    // self-energy slab is done (queued by stage_sigma), beside the latency-bound self-energy
    // pencil, traces and mu search.
    if (c->hf0_pending) {
        const double fact4 = c->p.U * (3.0 * c->Usp + c->Uch) / 4;
        HIPCHK(c, hipStreamWaitEvent(c->prep, c->ev_hf0, 0));
        HIPCHK(c, hipStreamWaitEvent(c->prep, c->ev_v, 0));
        c->hf0_pending = false;
        prof_begin(c, "jj_hf0_bnd", c->prep);
        HIPCHK(c, launch_hf0_bnd(c->g, c->d, c->mu0, fact4, c->prep));
        HIPCHK(c, hipEventRecord(c->ev_bnd, c->prep));
        prof_end(c);
        c->hv_slabs_due = true;
        c->hv_dropped = false;
        HIPCHK(c, hipEventRecord(c->ev_hv, c->prep));
        c->hv_pending = true;
    }
    return 0;
}

// the rest of the Lambda preparation (hV, f0V slabs and transposes) on the prep stream, after
// the work queued so far on the main stream
int queue_hv_slabs(tpsc_ctx* c)
{
    if (!c->hv_slabs_due) return 0;
    c->hv_slabs_due = false;
    HIPCHK(c, hipEventRecord(c->ev_sig, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->prep, c->ev_sig, 0));
    prof_begin(c, "jj_hVf0V_slab", c->prep);
    HIPCHK(c, launch_hVf0V_slab(c->g, c->d, c->prep));
    prof_end(c);
    prof_begin(c, "jj_transpose", c->prep);
    HIPCHK(c, launch_jj_transpose(c->g, c->d, c->prep));
    prof_end(c);
    HIPCHK(c, hipEventRecord(c->ev_hv, c->prep));
    return 0;
}

// green::calc_Self_FFT_spline slab + pencil parts, green.cpp:1737-2031
int stage_sigma(tpsc_ctx* c)
{
    if (!c->have_V) return fail(c, TPSC_E_ORDER, "self-energy before the vertex");
    const double fact = c->p.U * (3.0 * c->Usp + c->Uch) / 8.0;
    int rc = fork_side(c);
    if (rc) return rc;
    prof_begin(c, "sigma_bnd", c->side);
    HIPCHK(c, launch_sigma_bnd(c->g, c->d, c->mu0, fact, c->side));
    prof_end(c);
    prof_begin(c, "sigma_slab");
    HIPCHK(c, launch_sigma_slab(c->g, c->d, c->mu0, c->stream));
    prof_end(c);
    if ((rc = queue_hv_slabs(c))) return rc;
    if ((rc = join_side(c))) return rc;
    prof_begin(c, "sigma_pencil");
    HIPCHK(c, launch_sigma_pencil(c->g, c->d, c->stream));
    prof_end(c);
    // Sigma_inf is read by the host only for reports: its copy completes with the next wait
    HIPCHK(c, hipMemcpyAsync(c->host_scalar + 4, c->d.Sinf, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sinf_async = true;
    prof_flush(c);
    c->have_sigma = true;
    c->mu_ok = false;
    return 0;
}

// The trace reductions are reported values (nothing downstream reads them): each is queued
// into its own mapped slot and read at the next wait (traces_ready), so neither costs a host
// round trip of its own.
// green::traceSelfG, green.cpp:3001-3104: host_scalar[21], [22]
int stage_trace_selfG(tpsc_ctx* c)
{
    if (!c->have_sigma) return fail(c, TPSC_E_ORDER, "traceSelfG before the self-energy");
    HIPCHK(c, launch_trace_selfG(c->g, c->d, c->mu0, c->mu, 0, c->dev_scalar + 21, c->stream));
    HIPCHK(c, launch_trace_selfG(c->g, c->d, c->mu0, c->mu, 1, c->dev_scalar + 22, c->stream));
    c->trace_async |= 1;
    c->have_trSG = true;
    return 0;
}

// green::traceSelfG2, green.cpp:2929-2997: host_scalar[23]
int stage_trace_selfG2(tpsc_ctx* c)
{
    HIPCHK(c, launch_trace_selfG(c->g, c->d, c->mu0, c->mu, 2, c->dev_scalar + 23, c->stream));
    c->trace_async |= 2;
    c->have_trSG2 = true;
    return 0;
}

// the values of the trace reductions queued since the last read
int traces_ready(tpsc_ctx* c)
{
    if (!c->trace_async) return 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    volatile const double* h = c->host_scalar;
    if (c->trace_async & 1) {
        c->trSG = h[21];
        c->trSG_corr = h[21] + h[22];
    }
    if (c->trace_async & 2) {
        c->trSG2 = h[23];
        c->dblOcc2 = (h[23] + c->p.U * c->p.n * c->p.n / 4.0) / c->p.U;
    }
    c->trace_async = 0;
    return 0;
}

int eval_diffG(tpsc_ctx* c, double pmu, double* out)
{
    HIPCHK(c, launch_trace_diffG(c->g, c->d, pmu, c->mu0, c->dev_scalar, c->stream));
    return wait_mapped(c, out);
}

// green::find_mu, green.cpp:3185-3282
int stage_mu(tpsc_ctx* c)
{
    if (!c->have_sigma) return fail(c, TPSC_E_ORDER, "find_mu before the self-energy");
    int rc = 0;
    prof_begin(c, "mu_root");
    auto f = [&](double m) {
        double v = 0.0;
        if (rc == 0) rc = eval_diffG(c, m, &v);
        return v;
    };
    const double init[3] = {0.0, -10.0, 10.0}, lims[2] = {-20, 20};
    double root = 0.0;
    auto pass = [&]() { return launch_trace_diffG3_pass(c->g, c->d, c->dev_scalar + 8, c->stream); };
    int nh = 0;
    bool ok = false;
    rc = device_halley(c, kRootMu, pass, halley_start(init[0], lims[0], lims[1], 40, 0.0, 0.0), &root, &nh, &ok);
    if (rc) return rc;
    c->roots_dev = (c->roots_dev & ~4) | (ok ? 4 : 0);
    if (ok)
        c->nev_mu = nh;
    else
        ok = find_zero(f, init, &root, lims, &c->nev_mu);
    prof_end(c);
    prof_flush(c);
    if (rc) return rc;
    if (ok) {
        c->mu = root;
        c->mu_ok = true;
    }
    rc = stage_trace_selfG2(c);
    if (rc) return rc;
    if (!ok) return fail(c, TPSC_E_NOROOT, "mu not found");
    return 0;
}

// a device buffer of at least n doubles, kept by the context (grown when a call needs more)
int grow(tpsc_ctx* c, double** buf, size_t* cap, size_t n)
{
    if (*cap >= n) return 0;
    if (*buf) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(*buf));
        c->dev_bytes -= *cap * sizeof(double);
        *buf = nullptr;
        *cap = 0;
    }
    if (hipMalloc((void**)buf, n * sizeof(double)) != hipSuccess) {
        *buf = nullptr;
        return fail(c, TPSC_E_NOMEM, "device allocation failed");
    }
    *cap = n;
    c->dev_bytes += n * sizeof(double);
    return 0;
}

// the angle bin of every Lambda point (kx index l + 1, ky index m): floor(theta / (pi/(2 ntheta)))
// with theta = atan((N/2 - m)/(N/2 - l - 1)) (cond_opt.cpp:3174, 3792-3793), on the host
int angle_bins(tpsc_ctx* c, int ntheta)
{
    if (c->jjbins && c->jjbins_ntheta == ntheta) return 0;
    const Grid& g = c->g;
    if (!c->jjbins) {
        if (hipMalloc((void**)&c->jjbins, sizeof(int) * (size_t)std::max<long>(g.nq, 1)) != hipSuccess) {
            c->jjbins = nullptr;
            return fail(c, TPSC_E_NOMEM, "device allocation failed");
        }
        c->dev_bytes += sizeof(int) * (size_t)std::max<long>(g.nq, 1);
    }
    const int nk0 = g.N;
    const double Dtheta = kPi / (2.0 * ntheta);
    std::vector<int> b(g.nq);
    for (long q = 0; q < g.nq; q++) {
        const int l = (int)(q / g.nk), m = (int)(q % g.nk);
        const double th = atan((1.0 * (nk0 / 2 - m)) / (nk0 / 2 - l - 1));
        int it = (int)floor(th / Dtheta);
        b[q] = std::min(std::max(it, 0), ntheta - 1);
    }
    HIPCHK(c, hipMemcpy(c->jjbins, b.data(), sizeof(int) * g.nq, hipMemcpyHostToDevice));
    c->jjbins_ntheta = ntheta;
    return 0;
}

// cond_opt::calc_chijj_vertex_corr_all_optim (chi3 = false), cond_opt.cpp:2537-4290, with the
// contributions also binned by the angle of k in ntheta bins (cond_opt.cpp:3140-3174, 3792-3847)
int stage_chijj(tpsc_ctx* c, int ntheta)
{
    if (!c->have_sigma) {  // cond_opt.cpp:2550-2560
        int rc = tpsc_calc_self_energy_full(c);
        if (!rc) rc = stage_trace_selfG(c);
        if (!rc) rc = stage_mu(c);
        if (rc) return rc;
    }
    if (!c->mu_ok) return fail(c, TPSC_E_ORDER, "calc_chijj_vertex_corr_all_optim(): mu not computed");
    const Grid& g = c->g;
    const double fact4 = c->p.U * (3.0 * c->Usp + c->Uch) / 4;
    hipStream_t s = c->stream;
    // hV, f0V, their transposes and the boundary data were queued at the vertex stage (prep
    // stream); otherwise they are made here
    const bool prepared = c->hv_pending && !c->hv_dropped;
    c->hv_dropped = false;
    if (prepared) {
        // the kx = 0, pi rows and the pole data need no slab of the prep stream: they go on
        // the main stream while the hV/f0V slabs and transposes finish there (the poles read
        // the boundary kernel's hVb)
        prof_begin(c, "jj_kx0_rows");
        HIPCHK(c, launch_kx0_rows(g, c->d, c->mu, s));
        prof_end(c);
        HIPCHK(c, hipStreamWaitEvent(s, c->ev_bnd, 0));
        prof_begin(c, "jj_poles");
        HIPCHK(c, launch_jj_poles(g, c->d, c->mu, s));
        prof_end(c);
    }
    int rc = settle_prep(c, true);
    if (rc) return rc;
    if ((rc = fork_side(c))) return rc;
    if (!prepared) {
        prof_begin(c, "jj_hf0_bnd", c->side);
        HIPCHK(c, launch_hf0_bnd(g, c->d, c->mu0, fact4, c->side));
        prof_end(c);
        prof_begin(c, "jj_kx0_rows", c->side);
        HIPCHK(c, launch_kx0_rows(g, c->d, c->mu, c->side));
        prof_end(c);
        prof_begin(c, "jj_poles", c->side);
        HIPCHK(c, launch_jj_poles(g, c->d, c->mu, c->side));
        prof_end(c);
    }
    if (!prepared) {
        if (c->hf0_pending) {
            HIPCHK(c, hipStreamWaitEvent(s, c->ev_hf0, 0));
            c->hf0_pending = false;
        } else {
            prof_begin(c, "jj_hf0_slab");
            HIPCHK(c, launch_hf0_slab(g, c->d, s));
            prof_end(c);
        }
        prof_begin(c, "jj_hVf0V_slab");
        HIPCHK(c, launch_hVf0V_slab(g, c->d, s));
        prof_end(c);
        prof_begin(c, "jj_transpose");
        HIPCHK(c, launch_jj_transpose(g, c->d, s));
        prof_end(c);
    }
    const int nblk = (int)((g.nq + jj_points(g.Nw) - 1) / jj_points(g.Nw));
    const int nwb = g.nwb;
    double* part = c->d.jjpart;
    double* redrows = c->d.jjred;
    const int* bins = nullptr;
    if (ntheta > 1) {
        if ((rc = angle_bins(c, ntheta))) return rc;
        if ((rc = grow(c, &c->jjpart_t, &c->jjpart_t_cap, (size_t)nblk * 2 * nwb * ntheta))) return rc;
        if ((rc = grow(c, &c->jjred_t, &c->jjred_t_cap, (size_t)2 * nwb * ntheta))) return rc;
        part = c->jjpart_t;
        redrows = c->jjred_t;
        bins = c->jjbins;
    }
    if ((rc = join_side(c))) return rc;
    prof_begin(c, "jj_pencil");
    HIPCHK(c, launch_jj_pencil(g, c->d, c->mu, part, bins, ntheta, s));
    prof_end(c);
    HIPCHK(c, launch_reduce_cols2(part, 2 * nwb * ntheta, redrows, c->d.jjscal, 4, c->d.jjred + 2 * nwb, nblk,
                                  c->stream));
    std::vector<double> red(2 * nwb * ntheta), sc4(4), rows(2 * g.nk);
    HIPCHK(c, hipMemcpyAsync(red.data(), redrows, sizeof(double) * red.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(sc4.data(), c->d.jjred + 2 * nwb, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(rows.data(), c->d.kx0part, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_flush(c);
    // red: [b][2][nwb]; the totals add the bins in order
    c->ntheta = ntheta;
    c->chijj1.assign(nwb, 0.0);
    c->chijj2.assign(nwb, 0.0);
    c->chijj1_k.assign((size_t)nwb * ntheta, 0.0);
    c->chijj2_k.assign((size_t)nwb * ntheta, 0.0);
    for (int b = 0; b < ntheta; b++)
        for (int n = 0; n < nwb; n++) {
            const double v1 = red[(size_t)b * 2 * nwb + n], v2 = red[(size_t)b * 2 * nwb + nwb + n];
            c->chijj1_k[b + (size_t)n * ntheta] = v1;
            c->chijj2_k[b + (size_t)n * ntheta] = v2;
            c->chijj1[n] += v1;
            c->chijj2[n] += v2;
        }
    double k0 = sc4[0];
    const double i1 = sc4[1], i2 = sc4[2], i4 = sc4[3];
    for (double v : rows) k0 += v;
    c->kx0 = k0;
    c->jj_inf1 = i1;
    c->jj_inf2 = i2;
    c->jj_inf4 = i4;
    c->have_jj = true;
    return 0;
}

// sparse boson set of the Aslamazov-Larkin term (cond_opt.cpp:5123-5156): n = 0 ..
// (N0-1)/2^m, then m groups of (N0-1)/2^(m+1) indices with steps 2, 4, ..., 2^m; N0 = Nw/2+1
std::vector<long> al_freqs(int Nw, int mmin)
{
    const long N0 = Nw / 2 + 1, pw = 1L << mmin;
    const long jf0 = (N0 - 1) / pw, Nj = (N0 - 1) / (2 * pw);
    std::vector<long> f;
    for (long j = 0; j <= jf0; j++) f.push_back(j);
    long n = jf0, d = 1;
    for (int l = 0; l < mmin; l++) {
        d *= 2;
        for (long j = 0; j < Nj; j++) {
            n += d;
            f.push_back(n);
        }
    }
    return f;
}

int al_alloc(tpsc_ctx* c)
{
    const Grid& g = c->g;
    ALArrays& a = c->al;
    if (a.N == g.N && a.Nw == g.Nw) return 0;
    if (a.N) return fail(c, TPSC_E_INVALID, "Aslamazov-Larkin arrays were sized for another grid");
    const size_t plane = (size_t)(g.Nw + 1) * g.nq;
    int rc = 0;
    rc |= dalloc(c, &a.Grt, (size_t)(g.Nw + 1) * g.nk8);
    rc |= dalloc(c, &a.dGrt, 2 * (size_t)g.nk8);
    rc |= dalloc(c, &a.dhG, 2 * (size_t)g.nq);
    rc |= dalloc(c, &a.hw, (size_t)g.nq * (g.Nw / 2));
    rc |= dalloc(c, &a.hg2, (size_t)g.nq);
    rc |= dalloc(c, &a.chiwP, (size_t)g.nk8 * g.nwb);
    rc |= dalloc(c, &a.P1p, 2 * plane);
    rc |= dalloc(c, &a.P2p, 2 * plane);
    rc |= dalloc(c, &a.HGs, 2 * plane);
    rc |= dalloc(c, &a.P1s, 2 * plane);
    rc |= dalloc(c, &a.P2s, 2 * plane);
    rc |= dalloc(c, &a.dHq, 4 * (size_t)g.nq);
    rc |= dalloc(c, &a.dP, 8 * (size_t)g.nq);
    rc |= dalloc(c, &a.contrib, (size_t)g.nwb * g.nq);
    rc |= dalloc(c, &a.red, (size_t)g.nwb);
    if (rc) return rc;
    a.N = g.N;
    a.Nw = g.Nw;
    return 0;
}

// cond_opt.cpp:4292-6467 (chi3 = true): Lambda_jj3 on the sparse boson set, after the
// bubble and Maki-Thompson terms.  One launch sequence per boson frequency, no host sync
// until the fixed-order reduction of the per-k contributions at the end.
int stage_al(tpsc_ctx* c, int mmin, int ntheta)
{
    if (!c->have_jj) return fail(c, TPSC_E_ORDER, "Aslamazov-Larkin term needs Lambda_jj1, Lambda_jj2 first");
    if (mmin < 0 || mmin > 30) return fail(c, TPSC_E_INVALID, "mmin must be in 0..30");
    const Grid& g = c->g;
    int rc = al_alloc(c);
    if (rc) return rc;
    const std::vector<long> fr = al_freqs(g.Nw, mmin);
    const int Ns = (int)fr.size();
    hipStream_t s = c->stream;
    prof_begin(c, "al_setup");
    HIPCHK(c, launch_al_setup(g, c->d, c->al, c->mu0, s));
    prof_end(c);
    for (int n = 1; n < Ns; n++) {
        prof_begin(c, "al_freq");
        HIPCHK(c, launch_al_freq(g, c->d, c->al, c->Usp, c->Uch, c->mu, c->p.U, fr[n],
                                 c->al.contrib + (size_t)n * g.nq, s));
        prof_end(c);
    }
    std::vector<double> r(Ns, 0.0), rk((size_t)Ns * ntheta, 0.0);
    if (Ns > 1) {
        HIPCHK(c, launch_sum_rows(c->al.contrib + g.nq, Ns - 1, (int)g.nq, c->al.red + 1, s));
        HIPCHK(c, hipMemcpyAsync(r.data() + 1, c->al.red + 1, sizeof(double) * (Ns - 1), hipMemcpyDeviceToHost, s));
        if (ntheta > 1) {  // cond_opt.cpp:6261-6300, 6398
            if ((rc = angle_bins(c, ntheta))) return rc;
            if ((rc = grow(c, &c->al_red_t, &c->al_red_t_cap, (size_t)Ns * ntheta))) return rc;
            HIPCHK(c, launch_sum_rows_binned(c->al.contrib + g.nq, Ns - 1, (int)g.nq, c->jjbins, ntheta,
                                             c->al_red_t, s));
            HIPCHK(c, hipMemcpyAsync(rk.data() + ntheta, c->al_red_t, sizeof(double) * (size_t)(Ns - 1) * ntheta,
                                     hipMemcpyDeviceToHost, s));
        }
    }
    HIPCHK(c, hipStreamSynchronize(s));
    c->chijj3_k = (ntheta > 1) ? rk : std::vector<double>();
    prof_flush(c);
    c->al_freq = fr;
    c->al_mmin = mmin;
    c->chijj3 = r;
    c->chijj_all.assign(Ns, 0.0);
    c->chijj_all[0] = c->chijj1[0] + c->chijj2[0];
    for (int n = 1; n < Ns; n++) c->chijj_all[n] = c->chijj1[fr[n]] + c->chijj2[fr[n]] + r[n];
    return 0;
}

// ------------------------------------------------------------------ Sigma-loss gradient
int tan_alloc(tpsc_ctx* c)
{
    TanArrays& t = c->tan;
    if (t.chi) return 0;
    const Grid& g = c->g;
    const size_t nk8 = (size_t)g.nk8, nwb = (size_t)g.nwb;
    const size_t nblk = (size_t)((g.nk8 + 255) / 256) * kRowGroups;
    int rc = 0;
    rc |= dalloc(c, &t.dek, (size_t)g.nk * g.nk);
    rc |= dalloc(c, &t.dekw, nk8);
    rc |= dalloc(c, &t.der, nk8);
    rc |= dalloc(c, &t.chi, nwb * nk8);
    rc |= dalloc(c, &t.chirt, nwb * nk8);
    rc |= dalloc(c, &t.chirt0, nk8);
    rc |= dalloc(c, &t.dchirt0, nk8);
    rc |= dalloc(c, &t.dGrt0p, nk8);
    rc |= dalloc(c, &t.dchiqt0, nk8);
    rc |= dalloc(c, &t.chi0inf, nk8);
    rc |= dalloc(c, &t.chi0inf2, nk8);
    rc |= dalloc(c, &t.V, nwb * nk8);
    rc |= dalloc(c, &t.V0, nk8);
    rc |= dalloc(c, &t.dSig, 2 * nk8);
    rc |= dalloc(c, &t.Sb, 2 * nk8);
    rc |= dalloc(c, &t.Stau, (size_t)(g.Nw + 1) * nk8);
    rc |= dalloc(c, &t.Self, (size_t)(g.Nw / 2) * nk8);
    rc |= dalloc(c, &t.Sinf2, nk8);
    rc |= dalloc(c, &t.Sinf3, nk8);
    rc |= dalloc(c, &t.Sinf4, nk8);
    rc |= dalloc(c, &t.Sinf, 1);
    rc |= dalloc(c, &t.scr, 6 * nk8);
    rc |= dalloc(c, &t.part, 4 * nblk);
    rc |= dalloc(c, &t.red, 8);
    rc |= dalloc(c, &t.w, nk8);
    return rc;
}

// tangent of eps(k) and eps(r) for parameter p (0: t, 1: t', 2: mu0, 3: U) on the quadrant
// and the wedge; eps = -2t(cx+cy) - 4t' cx cy - 2t''(c2x+c2y) - 4t3(c2x cy + c2y cx),
// eps(r) = -t at r = (1,0), -t' at (1,1) (hamiltonien.cpp:22-29,119-122)
void tangent_model(const Grid& g, int p, std::vector<double>& dek, std::vector<double>& dekw, std::vector<double>& der)
{
    const int nk = g.nk;
    dek.assign((size_t)nk * nk, 0.0);
    dekw.assign((size_t)g.nk8, 0.0);
    der.assign((size_t)g.nk8, 0.0);
    if (p > 1) return;
    const double dk = M_PI / (nk - 1);
    for (int l = 0; l < nk; l++)
        for (int m = 0; m < nk; m++) {
            const double cx = cos(l * dk), cy = cos(m * dk);
            const double v = (p == 0) ? -2.0 * (cx + cy) : -4.0 * cx * cy;
            dek[(size_t)m + (size_t)l * nk] = v;
            if (m <= l) dekw[(size_t)m + (size_t)l * (l + 1) / 2] = v;
        }
    if (g.nk8 > 2) der[(p == 0) ? 1 : 2] = -1.0;
}

// tangent arrays of the Lambda_jj stage
int tan_alloc_jj(tpsc_ctx* c)
{
    TanArrays& t = c->tan;
    if (t.vel) return 0;
    const Grid& g = c->g;
    const size_t nq = (size_t)g.nq, plane = (size_t)(g.Nw + 1) * nq;
    t.jj_nblk = (int)std::min<long>(g.nq, 1024);
    int rc = 0;
    rc |= dalloc(c, &t.vel, nq);
    rc |= dalloc(c, &t.h, plane);
    rc |= dalloc(c, &t.f0, plane);
    rc |= dalloc(c, &t.hT, plane);
    rc |= dalloc(c, &t.f0T, plane);
    rc |= dalloc(c, &t.dhV, 2 * nq);
    rc |= dalloc(c, &t.df0V, 2 * nq);
    rc |= dalloc(c, &t.hVb, 2 * nq);
    rc |= dalloc(c, &t.f0Vb, 2 * nq);
    rc |= dalloc(c, &t.bscr, 12 * nq);
    rc |= dalloc(c, &t.jjpart, (size_t)t.jj_nblk * 2 * g.nwb);
    rc |= dalloc(c, &t.jjred, (size_t)2 * g.nwb + 2);
    rc |= dalloc(c, &t.jjtr, 2 * (size_t)((g.nk8 + 3) / 4));
    return rc;
}

// tangent of v_x(k) on the Lambda grid (kx index l + 1, ky index m) for parameter p:
// v_x = 2t sin kx + 4t' sin kx cos ky + ... (cond_opt.cpp:2645-2660 through k_model)
void tangent_velocity(const Grid& g, int p, std::vector<double>& dv)
{
    dv.assign((size_t)g.nq, 0.0);
    if (p > 1) return;
    const int nk = g.nk;
    const double dk = M_PI / (nk - 1);
    for (long i = 0; i < g.nq; i++) {
        const int l = (int)(i / nk), m = (int)(i % nk);
        const double sx = sin((l + 1) * dk), cy = cos(m * dk);
        dv[(size_t)i] = (p == 0) ? 2.0 * sx : 4.0 * sx * cy;
    }
}

// sum_k w_k Re Sigma(k, i w_0) (which 0) or Im Sigma (which 1) from the FP64 Sigma arrays given
int sigma_loss_value(tpsc_ctx* c, const double* Sinf, const double* Sinf2, const double* Sinf3, const cd* Self,
                     int which, double* out)
{
    HIPCHK(c, launch_sigma_loss(c->g, c->tan.w, Sinf, Sinf2, Sinf3, Self, which, c->tan.red, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, c->tan.red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// the Lambda_jj term of the loss: sum_{n <= 8} [Lambda_jj1 + Lambda_jj2](iq_n)
constexpr int kLossLambdaN = 8;
double lambda_term(const std::vector<double>& a, const std::vector<double>& b)
{
    double L = 0.0;
    for (size_t n = 0; n < a.size() && n <= (size_t)kLossLambdaN; n++) L += a[n] + b[n];
    return L;
}

// one tangent pass of the chain for parameter p; term[k] receives the derivative of loss term
// k (Re Sigma, Im Sigma, chi_sp(pi,pi), Lambda_jj) for the terms in c->loss_terms
int tangent_pass(tpsc_ctx* c, int p, double term[4])
{
    const Grid& g = c->g;
    TanArrays& t = c->tan;
    hipStream_t s = c->stream;
    const int terms = c->loss_terms;
    std::vector<double> dek, dekw, der;
    tangent_model(g, p, dek, dekw, der);
    HIPCHK(c, hipMemcpy(t.dek, dek.data(), sizeof(double) * dek.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.dekw, dekw.data(), sizeof(double) * dekw.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.der, der.data(), sizeof(double) * der.size(), hipMemcpyHostToDevice));
    const double dmu0 = (p == 2) ? 1.0 : 0.0, dU = (p == 3) ? 1.0 : 0.0;
    const double U = c->p.U, Usp = c->Usp, Uch = c->Uch, nsr = c->nsr;
    // chi0 -> tangent chi0(q,tau), chi0(r,tau) and tau = 0 data, then the (linear) pencil
    HIPCHK(c, launch_chi0_tan(g, c->d, t, c->mu0, dmu0, s));
    DevArrays v = c->d;
    v.chi = t.chi;
    v.chiw = t.chi;
    v.dchiqt0 = t.dchiqt0;
    v.chi0inf = t.chi0inf;
    v.chi0inf2 = t.chi0inf2;
    HIPCHK(c, launch_chi0_pencil(g, v, s));
    // implicit derivatives of the sum-rule roots (chi.cpp:833-843)
    double A[4];
    HIPCHK(c, launch_sum_rule_tan(g, c->d, t, Usp, Uch, t.red, s));
    HIPCHK(c, hipMemcpyAsync(A, t.red, sizeof A, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const double k2 = nsr * nsr / (2.0 * U);
    const double dUsp = (Usp * k2 / U * dU - A[0]) / (A[1] + k2);
    const double dUch = -(A[2] - k2 * dUsp + Usp * k2 / U * dU) / A[3];
    for (int k = 0; k < 4; k++) term[k] = 0.0;
    if (terms & TPSC_LOSS_CHISP_Q) {
        // chi_sp(Q, 0) = chi0/(1 - Usp chi0/2) at the last wedge point Q = (pi, pi), n = 0
        double x[2];
        HIPCHK(c, hipMemcpy(&x[0], c->d.chiw + (g.nk8 - 1), sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(&x[1], t.chi + (g.nk8 - 1), sizeof(double), hipMemcpyDeviceToHost));
        const double den = 1.0 - Usp * x[0] / 2.0;
        term[2] = (x[1] + x[0] * x[0] * dUsp / 2.0) / (den * den);
    }
    if (!(terms & (TPSC_LOSS_RE_SIGMA | TPSC_LOSS_IM_SIGMA | TPSC_LOSS_LAMBDA))) return 0;
    // V: tangent integrand, the linear tau transform, then + dfact chi0(r,tau)
    const double fact = U * (3.0 * Usp + Uch) / 8.0;
    const double dfact = (dU * (3.0 * Usp + Uch) + U * (3.0 * dUsp + dUch)) / 8.0;
    HIPCHK(c, launch_vertex_slab_tan(g, c->d, t, U, dU, Usp, dUsp, Uch, dUch, s));
    v = c->d;
    v.V = t.V;
    v.chirt = t.chirt;
    v.chirt0 = t.chirt0;
    v.V0 = t.V0;
    HIPCHK(c, launch_vertex_pencil(g, v, U, Usp, Uch, s));
    HIPCHK(c, launch_axpy((long)g.nwb * g.nk8, dfact, c->d.chirt, t.V, s));
    HIPCHK(c, launch_axpy(g.nk8, dfact, c->d.chirt0, t.V0, s));
    // Sigma: boundary data and slabs, then the (linear) pencil
    HIPCHK(c, launch_sigma_tan(g, c->d, t, c->mu0, dmu0, fact, dfact, s));
    v = c->d;
    v.Stau = t.Stau;
    v.Sb = t.Sb;
    v.dSig = t.dSig;
    v.Self = t.Self;
    v.Sinf2 = t.Sinf2;
    v.Sinf3 = t.Sinf3;
    v.Sinf4 = t.Sinf4;
    v.Sinf = t.Sinf;
    HIPCHK(c, launch_sigma_pencil(g, v, s));
    int rc = 0;
    if ((terms & TPSC_LOSS_RE_SIGMA) && (rc = sigma_loss_value(c, t.Sinf, t.Sinf2, t.Sinf3, t.Self, 0, &term[0])))
        return rc;
    if ((terms & TPSC_LOSS_IM_SIGMA) && (rc = sigma_loss_value(c, t.Sinf, t.Sinf2, t.Sinf3, t.Self, 1, &term[1])))
        return rc;
    if (!(terms & TPSC_LOSS_LAMBDA)) return 0;
    // mu from its root, Tr(G - G0)(mu) = 0 (green.cpp:3185-3282): dmu = -F_p / F_mu
    double F2[2];
    HIPCHK(c, launch_trace_diffG_tan(g, c->d, t, c->mu0, dmu0, c->mu, t.jjred, s));
    HIPCHK(c, hipMemcpyAsync(F2, t.jjred, sizeof F2, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const double dmu = -F2[0] / F2[1];
    // hV, f0V slabs and their boundary data, transposes, then the pencil
    std::vector<double> dv;
    tangent_velocity(g, p, dv);
    HIPCHK(c, hipMemcpy(t.vel, dv.data(), sizeof(double) * dv.size(), hipMemcpyHostToDevice));
    const double fact4 = 2.0 * fact, dfact4 = 2.0 * dfact;
    HIPCHK(c, launch_hVf0V_tan(g, c->d, t, c->mu0, dmu0, s));
    HIPCHK(c, launch_hf0_bnd_tan(g, c->d, t, c->mu0, dmu0, fact4, dfact4, s));
    v = c->d;
    v.h = t.h;
    v.f0 = t.f0;
    v.hT = t.hT;
    v.f0T = t.f0T;
    HIPCHK(c, launch_jj_transpose(g, v, s));
    HIPCHK(c, launch_jj_pencil_tan(g, c->d, t, c->mu, dmu, t.jj_nblk, s));
    const int nwb = g.nwb;
    HIPCHK(c, launch_reduce_cols(t.jjpart, t.jj_nblk, 2 * nwb, t.jjred, s));
    std::vector<double> red((size_t)2 * nwb);
    HIPCHK(c, hipMemcpyAsync(red.data(), t.jjred, sizeof(double) * red.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    std::vector<double> d1(red.begin(), red.begin() + nwb), d2(red.begin() + nwb, red.end());
    term[3] = lambda_term(d1, d2);
    c->grad_dmu[p] = dmu;
    for (int n = 0; n < nwb; n++) c->grad_dlambda[(size_t)p * nwb + n] = d1[n] + d2[n];
    return 0;
}

size_t field_count(tpsc_ctx* c, int field)
{
    const Grid& g = c->g;
    switch (field) {
        case TPSC_F_EK: return (size_t)g.nk * g.nk;
        case TPSC_F_CHI0:
        case TPSC_F_CHIRT:
        case TPSC_F_V:
        case TPSC_F_CHI0_TAU: return (size_t)g.nwb * g.nk8;
        case TPSC_F_CHI0_INF:
        case TPSC_F_CHI0_INF2:
        case TPSC_F_DTAU_CHIRT0:
        case TPSC_F_SIGMA_INF2:
        case TPSC_F_SIGMA_INF3:
        case TPSC_F_SIGMA_INF4: return (size_t)g.nk8;
        case TPSC_F_DTAU_SIGMA: return 2 * (size_t)g.nk8;
        case TPSC_F_SIGMA_BND: return 2 * (size_t)g.nk8;
        case TPSC_F_SIGMA_TAU: return (size_t)(g.Nw + 1) * g.nk8;
        case TPSC_F_SELF:
        case TPSC_F_SIGMA_KW: return 2 * (size_t)(g.Nw / 2) * g.nk8;
        case TPSC_F_CHIJJ1:
        case TPSC_F_CHIJJ2: return (size_t)g.nwb;
        case TPSC_F_CHIJJ3:
        case TPSC_F_CHIJJ_ALL:
        case TPSC_F_AL_FREQ: return c->chijj3.size();
        case TPSC_F_CHIJJ1_K:
        case TPSC_F_CHIJJ2_K: return c->chijj1_k.size();
        case TPSC_F_CHIJJ3_K: return c->chijj3_k.size();
        case TPSC_F_GRAD_TERMS: return c->grad_terms.size();
        case TPSC_F_GRAD_DMU: return c->grad_dmu.size();
        case TPSC_F_GRAD_DLAMBDA: return c->grad_dlambda.size();
        default: return 0;
    }
}

// The device stores Sigma(k,iw_n) point-major, Self[k*(Nw/2) + n]; copy-outs use the
// reference's frequency-major layout [n*nk8 + k] (green.cpp:2006).
// Remainder R(k,n) (FP64 path), complex128 [n*nk8 + k].
int get_self_remainder(tpsc_ctx* c, double* dst)
{
    const Grid& g = c->g;
    const int H = g.Nw / 2;
    std::vector<double> tmp(2 * (size_t)H * g.nk8);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->d.Self, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (long k = 0; k < g.nk8; k++)
        for (int j = 0; j < H; j++) {
            dst[2 * ((size_t)j * g.nk8 + k)] = tmp[2 * ((size_t)k * H + j)];
            dst[2 * ((size_t)j * g.nk8 + k) + 1] = tmp[2 * ((size_t)k * H + j) + 1];
        }
    return 0;
}

// Sigma(k,iw_n), n = 0..Nw/2-1, [n*nk8 + k] complex128: assembled from the remainder and the
// moments as in green.cpp:3039 (FP64 path) or widened from the stored FP32 values (mixed path)
int get_sigma_kw(tpsc_ctx* c, double* dst)
{
    if (!c->have_sigma) return fail(c, TPSC_E_ORDER, "self-energy not computed");
    if (int rc = sigma_inf_ready(c)) return rc;
    const Grid& g = c->g;
    const int H = g.Nw / 2;
    const size_t n = (size_t)H * g.nk8;
    hipStream_t s = c->stream;
    if (c->p.precision) {
        std::vector<float> tmp(2 * n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), c->d.Self, n * sizeof(cf), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        for (long k = 0; k < g.nk8; k++)
            for (int j = 0; j < H; j++) {
                dst[2 * ((size_t)j * g.nk8 + k)] = (double)tmp[2 * ((size_t)k * H + j)];
                dst[2 * ((size_t)j * g.nk8 + k) + 1] = (double)tmp[2 * ((size_t)k * H + j) + 1];
            }
        return 0;
    }
    int rc = get_self_remainder(c, dst);
    if (rc) return rc;
    std::vector<double> s2(g.nk8), s3(g.nk8);
    HIPCHK(c, hipMemcpyAsync(s2.data(), c->d.Sinf2, g.nk8 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(s3.data(), c->d.Sinf3, g.nk8 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    for (int j = 0; j < H; j++) {
        const double wn = (2 * j + 1) * kPi * g.T, w2 = wn * wn, w3 = w2 * wn, w4 = w2 * w2;
        for (long k = 0; k < g.nk8; k++) {
            double* z = dst + 2 * ((size_t)j * g.nk8 + k);
            const double re = -s2[k] / w2 + z[0] / w4;
            const double im = -c->Sigma_inf / wn + s3[k] / w3 + z[1] / w4;
            z[0] = re;
            z[1] = im;
        }
    }
    return 0;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

int tpsc_create(const tpsc_params* p, tpsc_ctx** out)
{
    if (!p || !out) return TPSC_E_INVALID;
    *out = nullptr;
    tpsc_ctx* c = new tpsc_ctx();
    c->p = *p;
    if (p->precision != TPSC_PREC_FP64 && p->precision != TPSC_PREC_MIXED) {
        *out = c;
        return fail(c, TPSC_E_INVALID, "precision must be TPSC_PREC_FP64 (0) or TPSC_PREC_MIXED (1)");
    }
    if (!supported_N(p->nk0) || !supported_Nw(p->Nw0) ||
        (p->coarse_pass && (p->nk0 < 8 || p->Nw0 < 32 || !supported_N(p->nk0 / 2)))) {
        *out = c;
        return fail(c, TPSC_E_INVALID, "unsupported grid");
    }
    if (!(p->T > 0.0)) {
        *out = c;
        return fail(c, TPSC_E_INVALID, "temperature must be positive");
    }
    if (!(p->n > 0.0 && p->n < 2.0)) {  // the density equation has a root only for 0 < n < 2
        *out = c;
        return fail(c, TPSC_E_INVALID, "density n must lie in (0, 2)");
    }
    *out = c;
    c->nsr = (p->n > 1.0) ? 2.0 - p->n : p->n;  // chi.cpp:18-19
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
        return fail(c, TPSC_E_DEVICE, "no HIP device available");
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIPCHK(c, hipStreamCreateWithFlags(&c->prep, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_hv, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_v, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_sig, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_tau0, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_bnd, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_hf0, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_mu0, hipEventDisableTiming));
    HIPCHK(c, hipHostMalloc((void**)&c->host_scalar, kHostScalars * sizeof(double), hipHostMallocMapped));
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->dev_scalar, c->host_scalar, 0));
    const int N = p->nk0, Nw = p->Nw0, nk = N / 2 + 1;
    const long nk8 = (long)nk * (nk + 1) / 2, nq = (long)nk * (nk - 2);
    const long nwb = Nw / 2 + 1;
    DevArrays& d = c->d;
    const size_t es = elem_bytes(c);  // tau / iqn arrays
    const size_t ss = p->precision ? sizeof(cf) : sizeof(cd);  // stored Sigma(k,iwn)
    int rc = 0;
    rc |= dalloc(c, &d.ek, (size_t)nk * nk);
    rc |= dalloc(c, &d.ekw, nk8);
    rc |= dalloc(c, &d.wq, nk8);
    rc |= dalloc(c, &d.er, std::max<long>(nk8, 5));
    rc |= dalloc(c, &d.chi, (size_t)nwb * nk8, es);
    if (p->precision)
        rc |= dalloc(c, &d.chiw, (size_t)nwb * nk8);
    else
        d.chiw = d.chi;
    rc |= dalloc(c, &d.chirt, (size_t)nwb * nk8, es);
    rc |= dalloc(c, &d.dchirt0, nk8);
    rc |= dalloc(c, &d.dGrt0p, nk8);
    rc |= dalloc(c, &d.dchiqt0, nk8);
    rc |= dalloc(c, &d.chi0inf, nk8);
    rc |= dalloc(c, &d.chi0inf2, nk8);
    rc |= dalloc(c, &d.srm, (size_t)(Nw + 1) * kSRM);
    rc |= dalloc(c, &d.srmax, (size_t)(Nw + 1));
    rc |= dalloc(c, &d.V, (size_t)nwb * nk8, es);
    rc |= dalloc(c, &d.Stau, (size_t)(Nw + 1) * nk8, es);
    rc |= dalloc(c, &d.dSig, 2 * nk8);
    rc |= dalloc(c, &d.Sb, 2 * nk8);
    rc |= dalloc(c, &d.chirt0, nk8);
    rc |= dalloc(c, &d.V0, nk8);
    rc |= dalloc(c, &d.Self, (size_t)(Nw / 2) * nk8, ss);
    rc |= dalloc(c, &d.Sinf2, nk8);
    rc |= dalloc(c, &d.Sinf3, nk8);
    rc |= dalloc(c, &d.Sinf4, nk8);
    rc |= dalloc(c, &d.Sinf, 1);
    rc |= dalloc(c, &d.dek, std::max<long>(nq, 1));
    rc |= dalloc(c, &d.d2ek, (size_t)nk * nk);
    rc |= dalloc(c, &d.h, (size_t)(Nw + 1) * std::max<long>(nq, 1), es);
    rc |= dalloc(c, &d.f0, (size_t)(Nw + 1) * std::max<long>(nq, 1), es);
    rc |= dalloc(c, &d.dhV, 2 * std::max<long>(nq, 1));
    rc |= dalloc(c, &d.df0V, 2 * std::max<long>(nq, 1));
    rc |= dalloc(c, &d.hVb, 2 * std::max<long>(nq, 1));
    rc |= dalloc(c, &d.f0Vb, 2 * std::max<long>(nq, 1));
    rc |= dalloc(c, &d.bscr, 4 * std::max<long>(nq, 1));
    rc |= dalloc(c, &d.hT, (size_t)(Nw + 1) * std::max<long>(nq, 1), es);
    rc |= dalloc(c, &d.f0T, (size_t)(Nw + 1) * std::max<long>(nq, 1), es);
    rc |= dalloc(c, &d.poles, 64 * (size_t)std::max<long>(nq, 1));
    const long nblk = (nq + jj_points(Nw) - 1) / jj_points(Nw);
    rc |= dalloc(c, &d.jjpart, (size_t)std::max<long>(nblk, 1) * 2 * nwb);
    rc |= dalloc(c, &d.jjscal, (size_t)std::max<long>(nblk, 1) * 4);
    rc |= dalloc(c, &d.kx0part, 2 * (size_t)nk);
    rc |= dalloc(c, &d.tmp, 4 * (size_t)nk8);
    rc |= dalloc(c, &d.part, kPartCap);
    rc |= dalloc(c, &d.jjred, 2 * (size_t)nwb + 4);
    rc |= dalloc(c, &d.red, 16);
    rc |= dalloc(c, &d.root, 1);
    rc |= dalloc(c, &d.twN, (size_t)Nw);
    rc |= dalloc(c, &d.twX, (size_t)Nw);
    rc |= dalloc(c, &d.twS, (size_t)Nw);
    rc |= dalloc(c, &d.twM, (size_t)Nw);
    rc |= dalloc(c, &d.ph, (size_t)Nw + 1);
    rc |= dalloc(c, &d.imax, 4);
    rc |= dalloc(c, &d.mu0v, 1);
    rc |= dalloc(c, &d.psc, kPscDoubles);
    rc |= dalloc(c, &d.ticket, 1);
    if (rc) return TPSC_E_NOMEM;
    HIPCHK(c, hipMemsetAsync(d.ticket, 0, sizeof(unsigned), c->stream));
    // = c->mu0 before any search.  On the context's stream: the streams are non-blocking, so a
    // memset on the null stream could land after the mu0 search of set_grid has written mu0v
    // (seen as a batch point solved at mu0 = 0 while other host threads kept the null stream busy)
    HIPCHK(c, hipMemsetAsync(d.mu0v, 0, sizeof(double), c->stream));
    return set_grid(c, N, Nw);  // hamiltonien + green0 constructors
}

void tpsc_destroy(tpsc_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->prep) (void)hipStreamSynchronize(c->prep);
    for (auto& pp : c->pending) {
        (void)hipEventDestroy(pp.a);
        (void)hipEventDestroy(pp.b);
    }
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->chi_tau) (void)hipFree(c->chi_tau);
    if (c->jjbins) (void)hipFree(c->jjbins);
    if (c->jjpart_t) (void)hipFree(c->jjpart_t);
    if (c->jjred_t) (void)hipFree(c->jjred_t);
    if (c->al_red_t) (void)hipFree(c->al_red_t);
    if (c->host_scalar) (void)hipHostFree(c->host_scalar);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_tau0) (void)hipEventDestroy(c->ev_tau0);
    if (c->ev_bnd) (void)hipEventDestroy(c->ev_bnd);
    if (c->ev_hf0) (void)hipEventDestroy(c->ev_hf0);
    if (c->ev_mu0) (void)hipEventDestroy(c->ev_mu0);
    for (auto& rg : c->root_graphs)
        if (rg.exec) (void)hipGraphExecDestroy(rg.exec);
    c->root_graphs.clear();
    if (c->ev_hv) (void)hipEventDestroy(c->ev_hv);
    if (c->ev_v) (void)hipEventDestroy(c->ev_v);
    if (c->ev_sig) (void)hipEventDestroy(c->ev_sig);
    if (c->prep) (void)hipStreamDestroy(c->prep);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* tpsc_last_error(const tpsc_ctx* c)
{
    if (c) return c->err.c_str();
    const char* m = tpsc_noalloc_error();
    return *m ? m : "null context";
}

int tpsc_find_mu0(tpsc_ctx* c, double* mu0)
{
    if (!c) return TPSC_E_INVALID;
    int rc = find_mu0(c);
    if (!rc) rc = mu0_ready(c);
    if (mu0) *mu0 = c->mu0;
    return rc;
}

int tpsc_chi0(tpsc_ctx* c, int nq0, int nw0)
{
    if (!c) return TPSC_E_INVALID;
    return stage_chi0(c, nq0, nw0);
}

int tpsc_sum_rule(tpsc_ctx* c, int channel, double Ux, double* value)
{
    if (!c || !value || (channel != 0 && channel != 1)) return TPSC_E_INVALID;
    return eval_sum_rule(c, channel, Ux, value);
}

int tpsc_find_Usp_Uch_grid(tpsc_ctx* c, int nq0, int nw0, double* Usp, double* Uch)
{
    if (!c) return TPSC_E_INVALID;
    double Usc[2] = {0.0, 0.0};
    bool ok = true;
    int rc = stage_UspUch_grid(c, nq0, nw0, Usc, &ok);
    if (Usp) *Usp = Usc[0];
    if (Uch) *Uch = Usc[1];
    if (rc) return rc;
    return ok ? 0 : fail(c, TPSC_E_NOROOT, "Usp or Uch root not found");
}

int tpsc_find_Usp_Uch(tpsc_ctx* c, double* Usp, double* Uch)
{
    if (!c) return TPSC_E_INVALID;
    int rc = stage_UspUch(c);
    if (Usp) *Usp = c->Usp;
    if (Uch) *Uch = c->Uch;
    return rc;
}

int tpsc_vertex_rt(tpsc_ctx* c)
{
    if (!c) return TPSC_E_INVALID;
    return stage_vertex(c);
}

int tpsc_self_energy(tpsc_ctx* c)
{
    if (!c) return TPSC_E_INVALID;
    return stage_sigma(c);
}

int tpsc_calc_self_energy_full(tpsc_ctx* c)
{
    if (!c) return TPSC_E_INVALID;
    int rc = stage_UspUch(c);  // green.cpp:1677-1678
    if (rc) return rc;
    rc = stage_vertex(c);      // green.cpp:1682
    if (rc) return rc;
    return stage_sigma(c);
}

int tpsc_trace_self_G(tpsc_ctx* c, double* trace, double* trace_corr)
{
    if (!c) return TPSC_E_INVALID;
    int rc = stage_trace_selfG(c);
    if (!rc) rc = traces_ready(c);
    if (trace) *trace = c->trSG;
    if (trace_corr) *trace_corr = c->trSG_corr;
    return rc;
}

int tpsc_trace_diff_G(tpsc_ctx* c, double mu, double* value)
{
    if (!c || !value) return TPSC_E_INVALID;
    if (!c->have_sigma) return fail(c, TPSC_E_ORDER, "traceDiffG before the self-energy");
    return eval_diffG(c, mu, value);
}

int tpsc_find_mu(tpsc_ctx* c, double* mu)
{
    if (!c) return TPSC_E_INVALID;
    int rc = stage_mu(c);
    if (mu) *mu = c->mu;
    return rc;
}

int tpsc_chijj(tpsc_ctx* c, int chi3, int mmin, int N_theta_k)
{
    if (!c) return TPSC_E_INVALID;
    if (N_theta_k < 1 || N_theta_k > kMaxTheta)
        return fail(c, TPSC_E_INVALID, "N_theta_k must be in 1.." + std::to_string(kMaxTheta));
    if (chi3 && (mmin < 0 || mmin > 30)) return fail(c, TPSC_E_INVALID, "mmin must be in 0..30");
    if (chi3 && c->p.Nw0 > 2048)
        return fail(c, TPSC_E_INVALID, "the Aslamazov-Larkin term needs Nw <= 2048 (its pencils hold 4 tau rows in LDS)");
    int rc = stage_chijj(c, N_theta_k);
    if (rc || !chi3) return rc;
    return stage_al(c, mmin, N_theta_k);
}

int tpsc_solve(tpsc_ctx* c)
{
    if (!c) return TPSC_E_INVALID;
    int rc = tpsc_calc_self_energy_full(c);
    if (!rc) rc = stage_trace_selfG(c);
    if (!rc) rc = stage_mu(c);
    if (!rc) rc = tpsc_chijj(c, 0, 0, 1);
    return rc;
}

size_t tpsc_field_size(tpsc_ctx* c, int field) { return c ? field_count(c, field) * sizeof(double) : 0; }

int tpsc_get(tpsc_ctx* c, int field, void* dst, size_t bytes)
{
    if (!c || !dst) return TPSC_E_INVALID;
    const size_t n = field_count(c, field);
    if (field == TPSC_F_CHIJJ3 || field == TPSC_F_CHIJJ_ALL || field == TPSC_F_AL_FREQ) {
        if (n == 0) return fail(c, TPSC_E_ORDER, "Aslamazov-Larkin term not computed");
        if (bytes < n * sizeof(double)) return fail(c, TPSC_E_INVALID, "destination too small");
        double* o = (double*)dst;
        for (size_t i = 0; i < n; i++)
            o[i] = (field == TPSC_F_CHIJJ3) ? c->chijj3[i] : (field == TPSC_F_CHIJJ_ALL) ? c->chijj_all[i]
                                                                                             : (double)c->al_freq[i];
        return 0;
    }
    if (field == TPSC_F_GRAD_TERMS || field == TPSC_F_GRAD_DMU || field == TPSC_F_GRAD_DLAMBDA) {
        if (n == 0) return fail(c, TPSC_E_ORDER, "no gradient of a loss with these terms computed");
        if (bytes < n * sizeof(double)) return fail(c, TPSC_E_INVALID, "destination too small");
        const std::vector<double>& v = (field == TPSC_F_GRAD_TERMS) ? c->grad_terms
                                       : (field == TPSC_F_GRAD_DMU) ? c->grad_dmu
                                                                    : c->grad_dlambda;
        memcpy(dst, v.data(), n * sizeof(double));
        return 0;
    }
    if (field == TPSC_F_CHIJJ1_K || field == TPSC_F_CHIJJ2_K || field == TPSC_F_CHIJJ3_K) {
        if (n == 0)
            return fail(c, TPSC_E_ORDER, field == TPSC_F_CHIJJ3_K ? "angle-binned Lambda_jj3 not computed (N_theta_k > 1)"
                                                                  : "Lambda_jj not computed");
        if (bytes < n * sizeof(double)) return fail(c, TPSC_E_INVALID, "destination too small");
        const std::vector<double>& v =
            (field == TPSC_F_CHIJJ1_K) ? c->chijj1_k : (field == TPSC_F_CHIJJ2_K) ? c->chijj2_k : c->chijj3_k;
        memcpy(dst, v.data(), n * sizeof(double));
        return 0;
    }
    if (n == 0) return fail(c, TPSC_E_INVALID, "unknown field");
    if (bytes < n * sizeof(double)) return fail(c, TPSC_E_INVALID, "destination too small");
    const void* src = nullptr;
    bool stored = false;  // array held in the tau / iqn storage precision
    switch (field) {
        case TPSC_F_EK: src = c->d.ek; break;
        case TPSC_F_CHIJJ1:
        case TPSC_F_CHIJJ2: {
            const std::vector<double>& v = (field == TPSC_F_CHIJJ1) ? c->chijj1 : c->chijj2;
            if (v.size() != n) return fail(c, TPSC_E_ORDER, "Lambda_jj not computed");
            memcpy(dst, v.data(), n * sizeof(double));
            return 0;
        }
        case TPSC_F_CHI0: src = c->d.chiw; break;
        case TPSC_F_CHIRT: src = c->d.chirt; stored = true; break;
        case TPSC_F_CHI0_INF: src = c->d.chi0inf; break;
        case TPSC_F_CHI0_INF2: src = c->d.chi0inf2; break;
        case TPSC_F_DTAU_CHIRT0: src = c->d.dchirt0; break;
        case TPSC_F_V: src = c->d.V; stored = true; break;
        case TPSC_F_SIGMA_TAU: src = c->d.Stau; stored = true; break;
        case TPSC_F_SELF:
            if (c->p.precision)
                return fail(c, TPSC_E_INVALID,
                            "the mixed-precision path stores the assembled Sigma(k,iwn): read TPSC_F_SIGMA_KW");
            return get_self_remainder(c, (double*)dst);
        case TPSC_F_SIGMA_KW: return get_sigma_kw(c, (double*)dst);
        case TPSC_F_SIGMA_INF2: src = c->d.Sinf2; break;
        case TPSC_F_SIGMA_INF3: src = c->d.Sinf3; break;
        case TPSC_F_SIGMA_INF4: src = c->d.Sinf4; break;
        case TPSC_F_DTAU_SIGMA: src = c->d.dSig; break;
        case TPSC_F_SIGMA_BND: src = c->d.Sb; break;
        case TPSC_F_CHI0_TAU:
            if (!c->chi_tau) return fail(c, TPSC_E_ORDER, "chi0(q,tau) not kept");
            src = c->chi_tau;
            stored = !c->chi_tau_fp64;
            break;
        default: return fail(c, TPSC_E_INVALID, "unknown field");
    }
    if (stored && c->p.precision) {
        std::vector<float> tmp(n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double* o = (double*)dst;
        for (size_t i = 0; i < n; i++) o[i] = (double)tmp[i];
        return 0;
    }
    HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// ------------------------------------------------------------------ spin correlation length
// eps(k) and its gradient on the host (hamiltonien.cpp:119-131), with the hopping t
double host_dispk(const tpsc_ctx* c, double kx, double ky)
{
    const double t = hop_t(c), tp = c->p.tp, tpp = c->p.tpp, t3 = c->p.t3;
    return -2.0 * t * (cos(kx) + cos(ky)) - 4.0 * tp * cos(kx) * cos(ky) - 2.0 * tpp * (cos(2 * kx) + cos(2 * ky)) -
           4 * t3 * (cos(2 * kx) * cos(ky) + cos(2 * ky) * cos(kx));
}

// Fermi velocity at the hot spot: the non-interacting Fermi surface is traced along kx - ky =
// kx0 for 10001 values of kx0 in [0, pi] (green0::find_FS, green0.cpp:626-675), then the
// segment that crosses kx + ky = pi gives the point where |grad eps| is taken
// (chi::vF_hot_spots, chi.cpp:419-447)
double hot_spot_vF(const tpsc_ctx* c)
{
    const int nk = 10001;
    std::vector<double> FS(2 * (size_t)nk, -1.0);
    for (int j = 0; j < nk; j++) {
        const double kx0 = j * M_PI / (nk - 1);
        const double init[3] = {0.5 * (kx0 + M_PI), kx0, M_PI}, lims[2] = {kx0, M_PI};
        double root = 0.0;
        auto f = [&](double kx) { return host_dispk(c, kx, kx - kx0) - c->mu0; };
        if (!find_zero(f, init, &root, lims, nullptr)) break;
        FS[2 * j] = root;
        FS[2 * j + 1] = root - kx0;
    }
    double kx1 = FS[0], ky1 = FS[1], kx2 = FS[2], ky2 = FS[3];
    int j = 1;
    while ((kx1 + ky1 - M_PI) * (kx2 + ky2 - M_PI) > 0 && j < nk - 1) {
        j++;
        kx1 = kx2;
        ky1 = ky2;
        kx2 = FS[2 * j];
        ky2 = FS[2 * j + 1];
    }
    const double kx = (kx1 + kx2) / 2, ky = (ky1 + ky2) / 2;
    const double t = hop_t(c), tp = c->p.tp, tpp = c->p.tpp, t3 = c->p.t3;
    const double gx = 2.0 * t * sin(kx) + 4.0 * tp * sin(kx) * cos(ky) + 4.0 * tpp * sin(2 * kx) +
                      4 * t3 * (2.0 * sin(2 * kx) * cos(ky) + cos(2 * ky) * sin(kx));
    const double gy = 2.0 * t * sin(ky) + 4.0 * tp * cos(kx) * sin(ky) + 4.0 * tpp * sin(2 * ky) +
                      4 * t3 * (cos(2 * kx) * sin(ky) + 2.0 * sin(2 * ky) * cos(kx));
    return sqrt(gx * gx + gy * gy);
}

// chi::sp_corr_length (chi.cpp:451-640) on chi0(q, iq_0) (the packed wedge c0): the inverse
// half width at half maximum of chi_sp(q, 0) along y through qmax and along x, and the
// asymptotic xi = xi0 sqrt(Usp/(2/chi0max - Usp)) with xi0^2 from the discrete curvature of chi0
// at qmax.  out = {vF, xi_y, xi_x, xi_as_y, xi_as_x}.  Where the reference reads a value it
// never set (chi_sp(qmax_x, 0) already above half maximum) 0 is used.
void sp_corr_from(const double* c0, int nbq, int q0x, int q0y, double chi0max, double Usp, double* out)
{
    auto w = [&](int l, int m) { return c0[m + (l * (l + 1)) / 2]; };
    auto sp = [&](double x) { return x / (1.0 - Usp * x / 2.0); };
    const double chisp_max = chi0max / (1.0 - Usp * chi0max / 2.0);
    const double dq = M_PI / (nbq - 1);
    int l = q0x, m = 0;
    double s = sp(w(l, 0)), s2 = 0.0;
    while (s < chisp_max / 2.0 && m <= q0y) {
        m++;
        s2 = s;
        s = sp(w(l, m));
    }
    double qa = (m - 1) * dq, qm = q0y * dq;
    double qHM = dq * (chisp_max / 2.0 - s2) / (s - s2) + qa;
    const double xi_y = 1.0 / (qm - qHM);
    double xi_x;
    m = q0y;
    l = m;
    s = sp(w(l, m));
    if (s < chisp_max / 2.0) {
        while (s < chisp_max / 2.0 && l <= q0x) {
            l++;
            s2 = s;
            s = sp(w(l, m));
        }
        qa = (l - 1) * dq;
        qm = q0x * dq;
        qHM = dq * (chisp_max / 2.0 - s2) / (s - s2) + qa;
        xi_x = 1.0 / (qm - qHM);
    } else
        xi_x = xi_y;
    double as_x, as_y;
    const double r = sqrt(Usp / (2.0 / chi0max - Usp));
    if (q0x == nbq - 1 && q0y == nbq - 1) {
        const double d2 = 2.0 * (w(nbq - 1, nbq - 2) - chi0max) * (nbq - 1) * (nbq - 1) / (M_PI * M_PI);
        as_x = sqrt(-0.5 * d2 / chi0max) * r;
        as_y = as_x;
    } else {
        int x = q0x, y = q0y - 1;
        if (y < 0) y = q0y + 1;
        const double cl = w(x, y), cm0 = w(q0x, q0y);
        y = q0y + 1;
        if (y > x) {
            x = q0y + 1;
            y = q0x;
        }
        const double cr = w(x, y);
        double d2 = (cr - 2 * cm0 + cl) * (nbq - 1) * (nbq - 1) / (M_PI * M_PI);
        as_y = sqrt(-0.5 * d2 / chi0max) * r;
        x = q0x - 1;
        y = q0y;
        if (y > x) {
            x = q0y;
            y = q0x - 1;
            if (y < 0) y = q0x + 1;
        }
        const double cl2 = w(x, y);
        x = q0x + 1;
        if (x > nbq - 1) x = q0x - 1;
        const double cr2 = w(x, q0y);
        d2 = (cr2 - 2 * cm0 + cl2) * (nbq - 1) * (nbq - 1) / (M_PI * M_PI);
        as_x = sqrt(-0.5 * d2 / chi0max) * r;
    }
    out[1] = xi_y;
    out[2] = xi_x;
    out[3] = as_y;
    out[4] = as_x;
}

int tpsc_sp_corr_length(tpsc_ctx* c, double* out)
{
    if (!c || !out) return TPSC_E_INVALID;
    if (!c->have_U || !c->have_chi0 || !full_grid(c))
        return fail(c, TPSC_E_ORDER, "sp_corr_length(): Usp and chi0 on the full grid are needed");
    const Grid& g = c->g;
    std::vector<double> c0((size_t)g.nk8);
    HIPCHK(c, hipMemcpyAsync(c0.data(), c->d.chiw, sizeof(double) * g.nk8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = hot_spot_vF(c);
    sp_corr_from(c0.data(), g.nk, c->qmax[0], c->qmax[1], c->chi0max, c->Usp, out);
    return 0;
}

int tpsc_fix_mu0(tpsc_ctx* c, double mu0)
{
    if (!c) return TPSC_E_INVALID;
    // a search still in flight (queued without a wait) is taken first, so that its result
    // cannot replace the value set here when it is read later
    int rc = mu0_ready(c);
    if (rc && rc != TPSC_E_NOROOT) return rc;
    if ((rc = invalidate_stages(c))) return rc;
    if (std::isnan(mu0)) {
        c->mu0_fixed = false;
        rc = find_mu0(c);
        return rc ? rc : mu0_ready(c);
    }
    c->mu0_fixed = true;
    c->mu0_failed = false;
    c->mu0 = mu0;
    return 0;
}

int tpsc_loss_forward(tpsc_ctx* c, const double* w, int terms, double* loss)
{
    if (!c || !loss) return TPSC_E_INVALID;
    const int all = TPSC_LOSS_RE_SIGMA | TPSC_LOSS_IM_SIGMA | TPSC_LOSS_CHISP_Q | TPSC_LOSS_LAMBDA;
    if (terms == 0 || (terms & ~all)) return fail(c, TPSC_E_INVALID, "unknown loss terms");
    if ((terms & (TPSC_LOSS_RE_SIGMA | TPSC_LOSS_IM_SIGMA)) && !w)
        return fail(c, TPSC_E_INVALID, "the Sigma terms need the weights w");
    if (c->p.precision != TPSC_PREC_FP64) return fail(c, TPSC_E_INVALID, "the gradient runs in the FP64 path only");
    if ((terms & TPSC_LOSS_LAMBDA) && !grad_jj_supported(c->g))
        return fail(c, TPSC_E_INVALID, "the Lambda_jj term of the gradient needs Nw <= 1024 and N >= 8");
    c->grad_ready = false;
    int rc = tpsc_calc_self_energy_full(c);
    if (rc) return rc;
    if ((rc = tan_alloc(c))) return rc;
    if ((terms & TPSC_LOSS_LAMBDA) && (rc = tan_alloc_jj(c))) return rc;
    if (w) HIPCHK(c, hipMemcpy(c->tan.w, w, sizeof(double) * c->g.nk8, hipMemcpyHostToDevice));
    double L = 0.0, v = 0.0;
    if (terms & TPSC_LOSS_RE_SIGMA) {
        if ((rc = sigma_loss_value(c, c->d.Sinf, c->d.Sinf2, c->d.Sinf3, c->d.Self, 0, &v))) return rc;
        L += v;
    }
    if (terms & TPSC_LOSS_IM_SIGMA) {
        if ((rc = sigma_loss_value(c, c->d.Sinf, c->d.Sinf2, c->d.Sinf3, c->d.Self, 1, &v))) return rc;
        L += v;
    }
    if (terms & TPSC_LOSS_CHISP_Q) {
        double x = 0.0;
        HIPCHK(c, hipMemcpy(&x, c->d.chiw + (c->g.nk8 - 1), sizeof(double), hipMemcpyDeviceToHost));
        L += x / (1.0 - c->Usp * x / 2.0);
    }
    if (terms & TPSC_LOSS_LAMBDA) {
        if ((rc = tpsc_trace_self_G(c, nullptr, nullptr)) || (rc = tpsc_find_mu(c, nullptr)) || (rc = tpsc_chijj(c, 0, 0, 1)))
            return rc;
        L += lambda_term(c->chijj1, c->chijj2);
    }
    c->loss_terms = terms;
    c->grad_terms.assign(16, 0.0);
    c->grad_dmu.assign(4, 0.0);
    c->grad_dlambda.assign((terms & TPSC_LOSS_LAMBDA) ? (size_t)4 * c->g.nwb : 0, 0.0);
    c->grad_ready = true;
    *loss = L;
    return 0;
}

int tpsc_loss_backward(tpsc_ctx* c, double dloss, double* grad)
{
    if (!c || !grad) return TPSC_E_INVALID;
    if (!c->grad_ready || !c->have_sigma)
        return fail(c, TPSC_E_ORDER, "the loss backward needs its forward first");
    if ((c->loss_terms & TPSC_LOSS_LAMBDA) && !c->have_jj)
        return fail(c, TPSC_E_ORDER, "Lambda_jj changed since the loss forward");
    for (int p = 0; p < 4; p++) {
        double term[4];
        int rc = tangent_pass(c, p, term);
        if (rc) return rc;
        double dL = 0.0;
        for (int k = 0; k < 4; k++) {
            c->grad_terms[(size_t)p * 4 + k] = term[k];
            dL += term[k];
        }
        grad[p] = dloss * dL;
    }
    return 0;
}

int tpsc_sigma_loss_forward(tpsc_ctx* c, const double* w, double* loss)
{
    if (!c || !w || !loss) return TPSC_E_INVALID;
    return tpsc_loss_forward(c, w, TPSC_LOSS_RE_SIGMA, loss);
}

int tpsc_sigma_loss_backward(tpsc_ctx* c, double dloss, double* grad) { return tpsc_loss_backward(c, dloss, grad); }

int tpsc_find_zero(double (*f)(double, void*), void* user, const double init[3], const double* lims, double* root,
                   int* nevals)
{
    if (!f || !init || !root) return TPSC_E_INVALID;
    const bool ok = find_root([&](double x) { return f(x, user); }, init, root, lims, nevals);
    return ok ? TPSC_OK : TPSC_E_NOROOT;
}

int tpsc_random_weights(unsigned long long seed, size_t n, double* out)
{
    if (!out) return TPSC_E_INVALID;
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (size_t i = 0; i < n; i++) out[i] = u(gen);
    return 0;
}

int tpsc_gwn_local(tpsc_ctx* c, double* out)
{
    if (!c || !out) return TPSC_E_INVALID;
    if (!c->have_sigma || !c->mu_ok) return fail(c, TPSC_E_ORDER, "G_local(): mu not computed or no self-energy");
    const Grid& g = c->g;
    const int H = g.Nw / 2;
    if (!c->gloc_buf) {
        int rc = dalloc(c, &c->gloc_buf, (size_t)(kGlocChunks + 1) * 2 * H);
        if (rc) return rc;
    }
    double* res = c->gloc_buf + (size_t)kGlocChunks * 2 * H;
    HIPCHK(c, launch_gloc(g, c->d, c->mu, c->gloc_buf, res, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, res, sizeof(double) * 2 * H, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int tpsc_set_params(tpsc_ctx* c, const tpsc_params* p)
{
    if (!c || !p) return TPSC_E_INVALID;
    if (p->nk0 != c->p.nk0 || p->Nw0 != c->p.Nw0 || p->precision != c->p.precision)
        return fail(c, TPSC_E_INVALID, "tpsc_set_params: the grid and the precision must stay those of the context");
    if (!(p->T > 0.0) || !(p->n > 0.0 && p->n < 2.0) ||
        (p->coarse_pass && (p->nk0 < 8 || p->Nw0 < 32 || !supported_N(p->nk0 / 2))))
        return fail(c, TPSC_E_INVALID, "tpsc_set_params: invalid parameters");
    c->p = *p;
    c->nsr = (p->n > 1.0) ? 2.0 - p->n : p->n;  // chi.cpp:18-19
    c->mu0_fixed = false;
    c->chi0_fp64 = false;  // the new point's enhancement decides again
    // a mu0 search of the previous point still in flight is taken (and its outcome dropped)
    int rc = mu0_ready(c);
    if (rc && rc != TPSC_E_NOROOT) return rc;
    if ((rc = invalidate_stages(c))) return rc;
    return set_grid(c, p->nk0, p->Nw0);  // the constructors' eps(k) and mu0 for the new point
}

int tpsc_get_params(tpsc_ctx* c, tpsc_params* out)
{
    if (!c || !out) return TPSC_E_INVALID;
    *out = c->p;
    return 0;
}

int tpsc_get_scalar(tpsc_ctx* c, int which, double* v)
{
    if (!c || !v) return TPSC_E_INVALID;
    if (which == TPSC_S_SIGMA_INF)
        if (int rc = sigma_inf_ready(c)) return rc;
    if (int rc = traces_ready(c)) return rc;
    if (int rc = chi0_ready(c)) return rc;
    if (int rc = mu0_ready(c)) return rc;
    switch (which) {
        case TPSC_S_MU0: *v = c->mu0; break;
        case TPSC_S_CHI0MAX: *v = c->chi0max; break;
        case TPSC_S_QMAX_L: *v = c->qmax[0]; break;
        case TPSC_S_QMAX_M: *v = c->qmax[1]; break;
        case TPSC_S_USP: *v = c->Usp; break;
        case TPSC_S_UCH: *v = c->Uch; break;
        case TPSC_S_DBLOCC: *v = c->dblOcc; break;
        case TPSC_S_DIFREL_SP: *v = c->difRelsp; break;
        case TPSC_S_DIFREL_CH: *v = c->difRelch; break;
        case TPSC_S_SIGMA_INF: *v = c->Sigma_inf; break;
        case TPSC_S_TRSG: *v = c->trSG; break;
        case TPSC_S_TRSG_CORR: *v = c->trSG_corr; break;
        case TPSC_S_MU: *v = c->mu; break;
        case TPSC_S_TRSG2: *v = c->trSG2; break;
        case TPSC_S_DBLOCC2: *v = c->dblOcc2; break;
        case TPSC_S_KX0: *v = c->kx0; break;
        case TPSC_S_CHIJJ1_INF2: *v = c->jj_inf1; break;
        case TPSC_S_CHIJJ2_INF2: *v = c->jj_inf2; break;
        case TPSC_S_CHIJJ2_INF4: *v = c->jj_inf4; break;
        case TPSC_S_NEVAL_SP: *v = c->nev_sp; break;
        case TPSC_S_NEVAL_CH: *v = c->nev_ch; break;
        case TPSC_S_NEVAL_MU: *v = c->nev_mu; break;
        case TPSC_S_NK: *v = c->g.nk; break;
        case TPSC_S_NK8: *v = (double)c->g.nk8; break;
        case TPSC_S_NWB: *v = c->g.nwb; break;
        case TPSC_S_NQ: *v = (double)c->g.nq; break;
        case TPSC_S_AL_N: *v = (double)c->chijj3.size(); break;
        case TPSC_S_AL_MMIN: *v = c->al_mmin; break;
        case TPSC_S_STAGES:
            *v = (double)((c->have_chi0 ? TPSC_STAGE_CHI0 : 0) | (c->have_U ? TPSC_STAGE_USP_UCH : 0) |
                          (c->have_V ? TPSC_STAGE_VERTEX : 0) | (c->have_sigma ? TPSC_STAGE_SIGMA : 0) |
                          (c->have_trSG ? TPSC_STAGE_TRACE_SG : 0) | (c->mu_ok ? TPSC_STAGE_MU : 0) |
                          (c->have_trSG2 ? TPSC_STAGE_TRACE_SG2 : 0) | (c->have_jj ? TPSC_STAGE_CHIJJ : 0) |
                          (c->chijj3.empty() ? 0 : TPSC_STAGE_AL));
            break;
        case TPSC_S_N_THETA: *v = c->ntheta; break;
        case TPSC_S_CHI0_FP64: *v = (c->p.precision == TPSC_PREC_FP64 || c->chi0_fp64) ? 1.0 : 0.0; break;
        case TPSC_S_ROOTS_ON_DEVICE: *v = c->roots_dev; break;
        default: return fail(c, TPSC_E_INVALID, "unknown scalar");
    }
    return 0;
}

int tpsc_prof_enable(tpsc_ctx* c, int on)
{
    if (!c) return TPSC_E_INVALID;
    c->prof = on != 0;
    return 0;
}
int tpsc_prof_reset(tpsc_ctx* c)
{
    if (!c) return TPSC_E_INVALID;
    prof_flush(c);
    c->prof_acc.clear();
    c->prof_names.clear();
    return 0;
}
int tpsc_prof_count(tpsc_ctx* c) { return c ? (int)c->prof_names.size() : 0; }
int tpsc_prof_get(tpsc_ctx* c, int i, const char** name, double* total_ms, int* calls)
{
    if (!c || i < 0 || i >= (int)c->prof_names.size()) return TPSC_E_INVALID;
    const std::string& n = c->prof_names[i];
    const auto& v = c->prof_acc[n];
    if (name) *name = n.c_str();
    if (total_ms) *total_ms = v.first;
    if (calls) *calls = v.second;
    return 0;
}

int tpsc_device_bytes(tpsc_ctx* c, double* bytes)
{
    if (!c || !bytes) return TPSC_E_INVALID;
    *bytes = (double)c->dev_bytes;
    return 0;
}

int tpsc_keep_chi_tau(tpsc_ctx* c, int on)
{
    if (!c) return TPSC_E_INVALID;
    c->keep_chi_tau = on != 0;
    if (on && !c->chi_tau) {
        const size_t n = (size_t)(c->p.Nw0 / 2 + 1) * ((long)(c->p.nk0 / 2 + 1) * (c->p.nk0 / 2 + 2) / 2);
        HIPCHK(c, hipMalloc((void**)&c->chi_tau, n * elem_bytes(c)));
    }
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ queued solves (batches)
// A slot holds d.psc (kPscDoubles), then the reduced Lambda rows (2 nwb) and the pencil's four
// scalar sums, then the kx = 0, pi rows of the diamagnetic term (2 nk); the kernels that make
// them write them there through the slot's device address.
size_t tpsc_queue_slot_doubles(const tpsc_ctx* c)
{
    return kPscDoubles + 2 * (size_t)c->g.nwb + 4 + 2 * (size_t)c->g.nk;
}

int tpsc_queue_solve(tpsc_ctx* c, const tpsc_params* p, double* slot_dev, bool chi0_fp64, bool long_searches)
{
    if (!c || !p || !slot_dev) return TPSC_E_INVALID;
    if (p->nk0 != c->p.nk0 || p->Nw0 != c->p.Nw0 || p->precision != c->p.precision || p->coarse_pass ||
        !full_grid(c))
        return fail(c, TPSC_E_INVALID, "a queued solve keeps the context's grid and precision, without a coarse pass");
    if (!(p->T > 0.0) || !(p->n > 0.0 && p->n < 2.0)) return fail(c, TPSC_E_INVALID, "invalid parameters");
    // the context's stages describe the previous point no longer; its mu0 search (if any) is dropped
    int rc = invalidate_stages(c);
    if (rc) return rc;
    c->mu0_async = false;
    c->mu0_fixed = false;
    c->chi0_fp64 = chi0_fp64 && p->precision;
    c->p = *p;
    c->nsr = (p->n > 1.0) ? 2.0 - p->n : p->n;  // chi.cpp:18-19
    Grid& g = c->g;
    g.T = p->T;
    DevArrays& d = c->d;
    hipStream_t s = c->stream;
    double* psc = d.psc;
    const double U = p->U, n = p->n, nsr = c->nsr;
    // Halley passes per search (TPSC_QUEUE_MAX_PASSES caps them: fewer points settle)
    const char* cap_env = getenv("TPSC_QUEUE_MAX_PASSES");
    const int cap = cap_env ? std::max(1, atoi(cap_env)) : 1 << 20;
    const int* np = long_searches ? kQueuedPassesLong : kQueuedPasses;
    const int nsp = std::min(np[0], cap), nch = std::min(np[1], cap), nmu = std::min(np[2], cap);
    // eps(k), then mu0 (into d.mu0v, the search's outcome into psc), as set_grid does
    HIPCHK(c, launch_model(g, d, hop_t(c), p->tp, p->tpp, p->t3, s));
    const double preset = (n == 1.0 && p->tp == 0 && p->tpp == 0) ? 0.0 : std::nan("");
    HIPCHK(c, launch_mu0_root(g, d, n, preset, psc + kPscMu0, s));
    // chi0(q,iqn), chi0(r,tau) and chi0max (stage_chi0), in FP64 for a promoted mixed point
    Grid gx = g;
    DevArrays dx = d;
    if (c->chi0_fp64) {
        if (!c->chirt64 && (rc = dalloc(c, &c->chirt64, (size_t)g.nwb * g.nk8))) return rc;
        gx.prec = 0;
        dx.chi = d.chiw;
        dx.chirt = c->chirt64;
    }
    HIPCHK(c, launch_chi0_slab(gx, dx, s));
    HIPCHK(c, launch_chi0_tau0(gx, dx, s));
    HIPCHK(c, launch_chi0_pencil(gx, dx, s));
    if (c->chi0_fp64) HIPCHK(c, launch_narrow(c->chirt64, (float*)d.chirt, (long)g.nwb * g.nk8, s));
    HIPCHK(c, launch_sr_moments(g, d, s));
    HIPCHK(c, launch_chi0max(g, d, s, psc + kPscChi0max));
    // Usp and Uch (stage_UspUch_grid): the same starts, brackets and host-known parts, with Uch's
    // constant formed on the device from Usp; each search settles into its own mirror in psc
    HalleyState s0 = halley_start(U / 2.0, 0.0, 0.0, 40, n - n * n / 2.0 - nsr, nsr * nsr / (2.0 * U));
    s0.T = g.T;
    HIPCHK(c, launch_halley_init(d.root, s0, d.red, psc + kPscUsp, s));
    for (int k = 0; k < nsp; k++) HIPCHK(c, launch_sum_rule3_pass(g, d, -1.0, psc + kPscUsp, s));
    s0 = halley_start(0.0, 0.0, 1000.0 * U, 40, 0.0, 0.0);
    s0.T = g.T;
    HIPCHK(c, launch_halley_init(d.root, s0, nullptr, psc + kPscUch, s, nullptr, psc + kPscUsp + 1, n - n * n / 2.0,
                                 nsr, U));
    for (int k = 0; k < nch; k++) HIPCHK(c, launch_sum_rule3_pass(g, d, 1.0, psc + kPscUch, s));
    // vertex and self-energy (stage_vertex, stage_sigma)
    HIPCHK(c, launch_vertex_slab(g, d, U, 0.0, 0.0, s, psc));
    HIPCHK(c, launch_vertex_pencil(g, d, U, 0.0, 0.0, s, psc));
    HIPCHK(c, launch_sigma_bnd(g, d, 0.0, 0.0, s, d.mu0v, U, psc));
    HIPCHK(c, launch_sigma_slab(g, d, 0.0, s, d.mu0v));
    HIPCHK(c, launch_sigma_pencil(g, d, s));
    // mu (stage_mu): start 0 in the limits (-20, 20), mu0 from the device
    s0 = halley_start(0.0, -20.0, 20.0, 40, 0.0, 0.0);
    s0.T = g.T;
    HIPCHK(c, launch_halley_init(d.root, s0, nullptr, psc + kPscMu, s, d.mu0v));
    for (int k = 0; k < nmu; k++) HIPCHK(c, launch_trace_diffG3_pass(g, d, psc + kPscMu, s));
    // Lambda_jj bubble + Maki-Thompson and the diamagnetic rows (stage_chijj, N_theta_k = 1)
    HIPCHK(c, launch_hf0_slab(g, d, s));
    HIPCHK(c, launch_hf0_bnd(g, d, 0.0, 0.0, s, d.mu0v, U, psc));
    HIPCHK(c, launch_hVf0V_slab(g, d, s));
    HIPCHK(c, launch_jj_transpose(g, d, s));
    double* const slot_rows = slot_dev + kPscDoubles;  // reduced Lambda rows, scalar sums, kx0 rows
    HIPCHK(c, launch_kx0_rows(g, d, 0.0, s, psc, slot_rows + 2 * g.nwb + 4));
    HIPCHK(c, launch_jj_poles(g, d, 0.0, s, psc));
    const int nblk = (int)((g.nq + jj_points(g.Nw) - 1) / jj_points(g.Nw));
    HIPCHK(c, launch_jj_pencil(g, d, 0.0, d.jjpart, nullptr, 1, s, psc));
    // the column sums straight into the slot, and the point scalars with them (block 0)
    HIPCHK(c, launch_reduce_cols2(d.jjpart, 2 * g.nwb, slot_rows, d.jjscal, 4, slot_rows + 2 * g.nwb, nblk, s, psc,
                                  kPscDoubles, slot_dev));
    return 0;
}

int tpsc_queue_sync(tpsc_ctx* c)
{
    if (!c) return TPSC_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int tpsc_queue_result(const tpsc_params* p, const double* slot, bool chi0_fp64, tpsc_batch_result* r, double* chijj,
                      int ldc)
{
    const int nwb = p->Nw0 / 2 + 1, nk = p->nk0 / 2 + 1;
    const double* psc = slot;
    // a failed search (or mu0 without a root) goes to tpsc_solve's bracketed fallback; one still
    // running when its passes ran out, to a queue with longer searches
    for (const int k : {kPscUsp, kPscUch, kPscMu})
        if (psc[k] == kRootFailed) return TPSC_E_NOROOT;
    if (psc[kPscMu0] == 0.0) return TPSC_E_NOROOT;
    for (const int k : {kPscUsp, kPscUch, kPscMu})
        if (psc[k] != kRootConverged) return kQueueLonger;
    const double Usp = psc[kPscUsp + 1], chi0max = psc[kPscChi0max];
    if (p->precision && !chi0_fp64 && 1.0 / (1.0 - Usp * chi0max / 2.0) > kPromoteEnhancement) return kQueuePromote;
    const double* red = slot + kPscDoubles;
    const double* sc4 = red + 2 * nwb;
    const double* rows = sc4 + 4;
    const double U = p->U, nsr = (p->n > 1.0) ? 2.0 - p->n : p->n;
    r->status = TPSC_OK;
    r->mu0 = psc[kPscMu0 + 1];
    r->Usp = Usp;
    r->Uch = psc[kPscUch + 1];
    r->mu = psc[kPscMu + 1];
    r->dblOcc = (p->n > 1.0) ? Usp / U * nsr * nsr / 4.0 + 1.0 - nsr : Usp / U * nsr * nsr / 4.0;
    double k0 = sc4[0];
    for (int i = 0; i < 2 * nk; i++) k0 += rows[i];
    r->kx0 = k0;
    if (chijj)
        for (int i = 0; i < std::min(nwb, ldc); i++) chijj[i] = (0.0 + red[i]) + (0.0 + red[nwb + i]);
    return TPSC_OK;
}
