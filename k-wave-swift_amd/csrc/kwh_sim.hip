// kwh_sim.hip -- host side of libkwavehip: descriptor validation, operator setup, the step
// schedule of the k-space first-order loop, sources, sensors, outputs, and the C ABI.
//
// Mirrors kspaceFirstOrder1D/2D/3D (Sources/KWave/Solver/KSpaceFirstOrder.swift:203-761). The
// host owns no field data after setup: every array lives in HBM and every step is a fixed
// sequence of launches on one HIP stream (optionally replayed from a captured hipGraph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <sys/stat.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kwave_hip.h"
#include "kwh_kernels.h"
#include "kwh_setup.h"
#include "kwh_adjoint.h"
#include "kwh_as.h"
#include "kwh_batch.h"
#include "kwh_launch.h"
#include "kwh_regops.h"
#include "kwh_planeops.h"

using namespace kwh;

namespace {

thread_local std::string g_last_error;

struct KwhError : std::runtime_error {
    int code;
    KwhError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define KWH_REQUIRE(cond, code, msg)                         \
    do {                                                     \
        if (!(cond)) throw KwhError((code), std::string(msg)); \
    } while (0)

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? KWH_ENOMEM : KWH_EDEVICE;
        throw KwhError(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define HIPCHK(x) hip_check((x), #x)

// Device buffer (RAII).
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        release();
        if (b == 0) return;
        HIPCHK(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename U> U* as() const { return reinterpret_cast<U*>(p); }
};

int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}

FftPlan make_plan(int L) {
    FftPlan pl{};
    pl.L = L;
    pl.nstages = 0;
    int m = L;
    auto push = [&](int r) {
        if (pl.nstages >= kMaxStages) throw KwhError(KWH_EUNSUPPORTED, "FFT length has too many factors");
        pl.radix[pl.nstages++] = r;
        m /= r;
    };
    while (m % 16 == 0) push(16);
    while (m % 8 == 0) push(8);
    while (m % 4 == 0) push(4);
    while (m % 2 == 0) push(2);
    while (m % 3 == 0) push(3);
    while (m % 5 == 0) push(5);
    for (int f = 7; m > 1; f += 2)
        while (m % f == 0) push(f);
    return pl;
}

std::vector<double> fft_freq(int n, double d) {
    std::vector<double> r(n, 0.0);
    const double val = 1.0 / (n * d);
    const int cutoff = (n - 1) / 2 + 1;
    for (int i = 0; i < cutoff; ++i) r[i] = i * val;
    int j = -(n / 2);
    for (int i = cutoff; i < n; ++i) r[i] = (j++) * val;
    return r;
}

// pmlProfile (Sources/KWave/PML/PML.swift:17-43); axisym: the radial axis of the axisymmetric
// solver has its PML at the outer edge only (:36-38).
std::vector<double> pml_profile(int n, double dx, double dt, double c, int pml_size, double pml_alpha, bool staggered,
                                bool axisym = false) {
    const double a = pml_alpha * (c / dx);
    const double m = pml_size;
    std::vector<double> pml(n, 1.0);
    for (int k = 0; k < std::min(pml_size, n); ++k) {
        const double x = k + 1;
        double pl, pr;
        if (staggered) {
            pl = (m + 0.5 - x) / m;
            pr = (x + 0.5) / m;
        } else {
            pl = (m + 1 - x) / m;
            pr = x / m;
        }
        if (!axisym) pml[k] = std::exp(-a * dt / 2.0 * std::pow(pl, 4.0));
        pml[n - pml_size + k] = std::exp(-a * dt / 2.0 * std::pow(pr, 4.0));
    }
    return pml;
}

double db2neper_unit(double y) { return 100.0 * std::pow(1e-6 / (2.0 * M_PI), y) / (20.0 * std::log10(M_E)); }

int grid_blocks(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 65536);
}

// Device scratch of the allocation-free entry point kwh_sim_read_output_noalloc: output staging
// (series rows, pRms, iAvg) in chunks of sensors that fit it. One copy per device, held until the
// process ends; every call on a device shares it and keeps the last call's contents (each call
// writes every scratch byte it reads back, so nothing needs clearing), and calls that overlap on
// one device race on it. In host code `scratch` only names the symbol: kernels and copies get
// the per-device address resolved by resolve_scratch.
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Launch-schedule cache of the autotuner, keyed by device and problem (grid shape, precision,
// register-resident axes): in memory for the process and persisted as text, so a schedule is
// timed once per machine. File: KWH_AUTOTUNE_CACHE (a path, or "off"), else
// $XDG_CACHE_HOME/kwavehip/autotune.txt, else $HOME/.cache/kwavehip/autotune.txt. One line per
// entry: "<device> <problem> <ntr_shift> <p_shift> <lds bytes> <min_blocks>"; the device is the
// gfx architecture name and compute-unit count, since schedules do not carry across GPUs.
struct TunedSchedule {
    int ntr_shift = 0, p_shift = 0;
    size_t lds = 64 * 1024;
    int min_blocks = 1024;
};
std::mutex g_tune_mu;
std::map<std::string, TunedSchedule> g_tune;
std::map<std::string, bool> g_tune_files_read;

std::string tune_cache_path() {
    if (const char* e = std::getenv("KWH_AUTOTUNE_CACHE")) return std::strcmp(e, "off") == 0 ? std::string() : e;
    if (const char* x = std::getenv("XDG_CACHE_HOME")) return std::string(x) + "/kwavehip/autotune.txt";
    if (const char* h = std::getenv("HOME")) return std::string(h) + "/.cache/kwavehip/autotune.txt";
    return std::string();
}

// with g_tune_mu held: merge the file's entries into the in-memory cache (once per path)
void tune_read(const std::string& path) {
    if (path.empty() || g_tune_files_read[path]) return;
    g_tune_files_read[path] = true;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char dev[128], prob[128];
    TunedSchedule t;
    unsigned long lds = 0;
    while (std::fscanf(f, "%127s %127s %d %d %lu %d", dev, prob, &t.ntr_shift, &t.p_shift, &lds, &t.min_blocks) == 6) {
        t.lds = (size_t)lds;
        g_tune[std::string(dev) + " " + prob] = t;
    }
    std::fclose(f);
}

// with g_tune_mu held: append one entry (directories created as needed); false if not written
bool tune_write(const std::string& path, const std::string& key, const TunedSchedule& t) {
    if (path.empty()) return false;
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0755);
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return false;
    std::fprintf(f, "%s %d %d %lu %d\n", key.c_str(), t.ntr_shift, t.p_shift, (unsigned long)t.lds, t.min_blocks);
    std::fclose(f);
    return true;
}

std::string device_tag(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string arch(prop.gcnArchName);
    for (char& c : arch)
        if (c == ' ' || c == ':') c = '_';
    return arch + "-" + std::to_string(prop.multiProcessorCount) + "cu";
}

}  // namespace

struct kwh_ctx {
    int device;
};

// The launches of one step of one sim for one step key, with a hash of their kernels and
// fixed workgroup sizes: sims whose hashes match can share batched launches.
struct StepRecord {
    unsigned key = 0;
    uint64_t signature = 0;
    std::vector<LaunchRec> launches;
};

// Abstract sim: the typed implementation follows.
struct kwh_sim {
    virtual ~kwh_sim() = default;
    virtual void launch(int t0, int t1) = 0;
    virtual void sync() = 0;
    virtual void run(int t0, int t1, kwh_step_cb cb, int interval, void* user) = 0;
    virtual void output_shape(int field, int64_t shape[3], int* ndim) = 0;
    virtual void read_output(int field, void* dst, size_t bytes) = 0;
    virtual void read_output_noalloc(int field, void* dst, size_t bytes, void* scratch_ptr, size_t scratch_bytes) = 0;
    virtual int steps_done() const = 0;
    virtual size_t state_bytes() = 0;
    virtual void save_state(void* dst, size_t bytes) = 0;
    virtual void load_state(const void* src, size_t bytes) = 0;
    virtual void set_timing(bool on) = 0;
    virtual void set_timing_only(const char* name) = 0;
    virtual int kernel_stats(kwh_kernel_stat* st, int cap) = 0;
    virtual void set_graph(bool on) = 0;
    virtual void set_plane(bool on) = 0;
    virtual bool plane_mode() const = 0;
    virtual bool plane_available() const = 0;
    virtual void element_traces(int field, const int32_t* element, const int32_t* delay, int nel, double scale,
                                void* dst, size_t bytes) = 0;
    virtual void set_checkpointing(int every) = 0;
    virtual void set_nonfinite_check(int every) = 0;
    virtual std::string autotune() = 0;
    virtual void time_reversal_result(double factor, void* dst) = 0;
    virtual void vjp(const void* dldy, int every, void* grad_p0, void* grad_c0) = 0;
    // batched runs (kwh_batch)
    virtual int total_steps() const = 0;
    virtual void step_done(int t) = 0;
    virtual void set_counters_on(hipStream_t s) = 0;
    virtual void graph_step(int t) = 0;   // replay step t's graph on the sim's own stream
    virtual int device() const = 0;
    virtual const StepRecord& recorded_step(int t) = 0;   // step t's launches, recorded once per step key
    virtual uint64_t shape_class() const = 0;            // grid shape and precision
    virtual int elem_bytes() const = 0;                  // 4 (float32) or 8 (float64)
};

namespace {

constexpr int kNames = 23;
const char* kKernelNames[kNames] = {"strided_fwd", "strided_inv", "middle_grad", "middle_div", "middle_abs",
                                    "z_velocity",  "z_density",   "z_eos",       "z_init",     "z_forward_real",
                                    "z_inverse_real", "middle_table", "gather", "source_scatter", "fill", "other",
                                    "z_colloc_rows", "plane_velocity", "plane_density", "plane_forward_real",
                                    "plane_inverse_real", "plane_eos", "z_rhot"};
enum KId { K_SF, K_SI, K_MG, K_MD, K_MA, K_ZV, K_ZD, K_ZE, K_ZI, K_ZFR, K_ZIR, K_MT, K_GA, K_SC, K_FI, K_OT, K_ZCR,
           K_PV, K_PD, K_PFR, K_PIR, K_PE, K_ZRT };

template <typename T>
struct SourceDev {
    bool present = false;
    int mode = 0;
    int rows = 0, len = 0;
    int64_t n = 0;
    DBuf index, signal, base, rowptr;
    double factor[3] = {1, 1, 1};
};

template <typename T>
struct SimT final : kwh_sim {
    using C = cx<T>;
    int dev = 0;
    hipStream_t stream = nullptr;
    Geo g{};
    int dim = 0;
    int nphys[3] = {1, 1, 1};
    double dphys[3] = {1, 1, 1};
    double dt = 0;
    int nt = 0;
    int mid_axis = 1;
    bool plain1 = false;   // axis-1 plain passes (3D)
    int steps = 0;
    double c_ref = 0;

    // tables
    AxisTabs<T> tab{};
    AxisTabs<double> tab64{};
    DBuf tw_buf[3], ddxp_buf[3], ddxn_buf[3], coll_buf[3], pml_buf[3], pmlsg_buf[3];
    DBuf tw64_buf[3], kv64_buf[3];
    DBuf kappa, skappa, nab1, nab2;

    // media
    DBuf c0m, rho0m, dtrm[3], taum, etam, bonam;
    SMap<T> c0s{}, rho0s{}, dtrs[3]{}, taus{}, etas{}, bonas{};
    bool nonlinear = false, absorbing = false, stokes = false, dispersion = false;
    bool tau_eta_from_c = false;        // tau = tk c^te, eta = ek c^ee evaluated in the EOS kernel
    double tk = 0, te = 0, ek = 0, ee = 0;

    // state
    DBuf p, u[3], rho[3], rhoT, ucol[3], p0s;
    DBuf spec[6];          // spectra work buffers
    DBuf scr_real, scr_spec, corr_u[3], corr_p;
    // k-space-corrected sources: the middle-axis slabs [slab_lo, slab_lo + slab_n) that hold every
    // source point (slab_n = 0: not used). The correction's forward passes transform only these
    // slabs, into src_spec, whose other slabs stay zero from setup on.
    DBuf src_spec;
    int slab_lo = 0, slab_n = 0;
    bool has_p0 = false;
    bool P_valid = false;  // spec[0] holds R2C_z(p) (R2C_yz(p) in plane mode) for the next step

    // sources
    SourceDev<T> psrc, usrc[3];

    // sensor
    int64_t nsens = 0;
    int ncorner = 1;
    bool cart = false;
    DBuf sidx, sw;
    uint32_t record = 0;
    bool rec_p = false, rec_u = false, store_p = false, store_u = false;
    bool need_u[3] = {false, false, false};   // collocated component a is computed each step
    int ucol_slot[3] = {-1, -1, -1};          // spectrum buffer index holding its colloc spectrum
    // rows (flat / m2) read by the sensor; the contiguous-axis component is collocated on them only
    DBuf zrows;
    int nzrows = 0;
    DBuf ser_p, ser_u[3], pmax, pmin, psum2, umax[3], usum2[3];
    int rec_steps = 0;
    // device step counters {t, record slot}, advanced by a kernel at the end of every step
    DBuf ctr;
    // checkpoints of (p, u_a, rho_a) after every ckpt_every-th step, for the adjoint
    int ckpt_every = 0;
    bool smoothed_media = false, smoothed_p0 = false;   // the adjoint does not cover smoothing
    DBuf ckpt;
    std::vector<char> ckpt_have;
    // captured step graphs, keyed by step_key(t)
    struct StepGraph { hipGraph_t graph; hipGraphExec_t exec; };
    std::vector<std::pair<unsigned, StepGraph>> step_graphs;
    // directivity
    int ndir = 0;
    std::vector<std::unique_ptr<DBuf>> dir_w, dir_idx, dir_pos;
    std::vector<int64_t> dir_count;
    DBuf dir_field;

    // timing
    bool timing = false;
    int timing_only = -1;   // >= 0: only launches of this kernel id are timed
    struct Pending { int id; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    int64_t stat_launch[kNames] = {};
    double stat_ms[kNames] = {};
    double stat_bytes[kNames] = {};
    // graph replay of the time step
    bool use_graph = false;
    // optional non-finite check of p every nonfinite_every steps (0 = off); nf_flag holds the
    // smallest step index after which p held a NaN or an infinity (INT_MAX: none)
    int nonfinite_every = 0;
    DBuf nf_flag;

    // one-element device buffers behind scalar SMaps (branch-free at() in the kernels)
    DBuf scal_pool;
    int scal_n = 0;
    SMap<T> zero_map{};
    SMap<T> mk_map(const T* p) const { return SMap<T>{p, T(0), p, ~(int64_t)0}; }
    SMap<T> mk_scalar(T v) {
        KWH_REQUIRE(scal_n < 64, KWH_EUNSUPPORTED, "scalar pool exhausted");
        T* d = scal_pool.as<T>() + (size_t)kScalarRep * scal_n++;
        const std::vector<T> rep(kScalarRep, v);
        HIPCHK(hipMemcpy(d, rep.data(), kScalarRep * sizeof(T), hipMemcpyHostToDevice));
        return SMap<T>{nullptr, v, d, 0};
    }
    // register-resident launchers per internal axis (null: generic Stockham kernels)
    const RegOps<T>* rx[3] = {nullptr, nullptr, nullptr};
    // y-z plane kernels for this grid's (m1, m2) plane (null: none fits); plane mode runs the y
    // and z passes of a step inside them, so spec[0] carries R2C_yz(p) between steps
    const PlaneOps<T>* px = nullptr;
    bool plane = false;
    bool plane_wanted = true;   // on by default where plane mode applies (KWH_PLANE=0: off)
    // launch schedule chosen by the autotuner: halve the tile columns of the strided / middle
    // passes (ntr_shift), halve the row pairs per workgroup of the row kernels (p_shift), LDS
    // budget of the generic Stockham passes and the workgroup count below which they take fewer
    // transforms per workgroup (min_blocks; 1024 = four per CU)
    struct Schedule {
        int ntr_shift = 0, p_shift = 0;
        size_t lds = 64 * 1024;
        int min_blocks = 1024;
        // register strided passes walk each XCD's run of tiles from its end: the row kernels
        // around them walk theirs from the start, so each pass begins with the x planes the one
        // before touched last (still in the memory-side cache); 256^3 y passes 58.9 -> 55.5 us
        int strided_rev = 1;
        int grad_dual = 1;   // KWH_GRAD_DUAL=0: three x-pass outputs for the gradient
        int corr_fuse = 1;   // KWH_CORR_FUSE=0: the plane source correction as a real field
        int row_half = 1;    // KWH_ROW_HALF=0: length-512 rows as row pairs in every row kernel
    };
    Schedule sched;
    bool use_sq = true;

    // axisymmetric solver (kspaceFirstOrderAS, KSpaceFirstOrderAS.swift:17-309): the fields live
    // on the physical [nx][ny] grid (x axial, y radial); every spectral operation mirrors its
    // input into the 4x radially expanded grid [nx][4 ny] and transforms there. The expanded
    // geometry, tables and launchers form a second spectral space that ExScope swaps in while
    // the transform passes run.
    bool axisym = false;
    struct Space { Geo g; AxisTabs<T> tab; const RegOps<T>* rx[3]; int mid_axis; bool plain1; };
    Space ex{};
    DBuf ex_tw[3], ex_ddxp[3], ex_ddxn[3], ex_coll[3], ex_kv[3], ex_kappa, ex_skappa, invr_buf;
    DBuf ex_real[3], ex_spec[3];
    void swap_space() {
        std::swap(g, ex.g);
        std::swap(tab, ex.tab);
        for (int a = 0; a < 3; ++a) std::swap(rx[a], ex.rx[a]);
        std::swap(mid_axis, ex.mid_axis);
        std::swap(plain1, ex.plain1);
    }
    struct ExScope {
        SimT* s;
        explicit ExScope(SimT* owner) : s(owner) { s->swap_space(); }
        ~ExScope() { s->swap_space(); }
        ExScope(const ExScope&) = delete;
        ExScope& operator=(const ExScope&) = delete;
    };

    ~SimT() override {
        for (auto& sg : step_graphs) {
            if (sg.second.exec) (void)hipGraphExecDestroy(sg.second.exec);
            (void)hipGraphDestroy(sg.second.graph);
        }
        for (auto& e : ev_pool) (void)hipEventDestroy(e);
        for (auto& pd : pending) {
            (void)hipEventDestroy(pd.a);
            (void)hipEventDestroy(pd.b);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ------------------------------------------------------------------ timing helpers
    hipEvent_t get_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    template <typename F>
    void timed(int id, double bytes, F&& f) {
        if (!timing || recording || (timing_only >= 0 && id != timing_only)) {
            f();
            HIPCHK(hipGetLastError());
            return;
        }
        hipEvent_t a = get_event(), b = get_event();
        HIPCHK(hipEventRecord(a, stream));
        f();
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b, stream));
        pending.push_back({id, a, b, bytes});
    }
    void collect_timing() {
        for (auto& pd : pending) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, pd.a, pd.b));
            stat_launch[pd.id] += 1;
            stat_ms[pd.id] += ms;
            stat_bytes[pd.id] += pd.bytes;   // launches of one kernel may differ in size (fields per pass)
            ev_pool.push_back(pd.a);
            ev_pool.push_back(pd.b);
        }
        pending.clear();
    }

    // ------------------------------------------------------------------ launches
    // Every launch of the step goes through emit: launched on the sim's stream, or, while a
    // batched run records this sim's step, appended to `recording`.
    std::vector<LaunchRec>* recording = nullptr;
    template <class B, class... A>
    void emit(dim3 grid, dim3 block, size_t lds, const A&... a) {
        if (recording) {
            recording->push_back(make_rec<B>(grid, block, lds, a...));
            return;
        }
        hipLaunchKernelGGL(B::kernel, grid, block, lds, stream, a...);
    }
    Emit em() const { return Emit{stream, recording}; }

    // ------------------------------------------------------------------ launch geometry
    struct ZCfg { ZTile zt; int blocks, threads; size_t lds; };
    ZCfg zcfg(int nbuf) const {
        const int L = g.m2;
        const int npairs = (g.nrows + 1) / 2;
        int ntr = 32;
        while (ntr > 1 && ((size_t)nbuf * L * (ntr + 1) * sizeof(C) > sched.lds || ntr / 2 >= npairs)) ntr >>= 1;
        while (ntr > 1 && (npairs + ntr - 1) / ntr < sched.min_blocks) ntr >>= 1;   // fill the CUs first
        ZCfg c{};
        c.zt.ntr = ntr;
        c.zt.lg_ntr = ilog2(ntr);
        c.zt.pitch = ntr > 1 ? ntr + 1 : 1;
        c.zt.q0 = 0;
        c.blocks = (npairs + ntr - 1) / ntr;
        int thr = ((ntr * L / 8 + 63) / 64) * 64;
        c.threads = std::max(64, std::min(256, thr));
        c.lds = (size_t)nbuf * L * c.zt.pitch * sizeof(C);
        KWH_REQUIRE(c.lds <= 65536, KWH_EUNSUPPORTED, "contiguous axis too long for LDS staging");
        return c;
    }
    struct SCfg { int64_t stride, inner, ncols; int L, lg_ntr, pitch, blocks, threads; size_t lds; };
    SCfg scfg(int axis, int nbuf) const {
        SCfg c{};
        c.L = axis == 0 ? g.m0 : g.m1;
        if (axis == 0) { c.stride = (int64_t)g.m1 * g.mzh; c.inner = c.stride; c.ncols = c.stride; }
        else { c.stride = g.mzh; c.inner = g.mzh; c.ncols = (int64_t)g.m0 * g.mzh; }
        const int L = std::max(c.L, 1);
        int ntr = 64;
        while (ntr > 1 && ((size_t)nbuf * L * ntr * sizeof(C) + 64 * sizeof(int64_t) > sched.lds || ntr / 2 >= c.ncols))
            ntr >>= 1;
        // fill the CUs first, keeping at least 8 adjacent columns (64 B rows in float32) per tile
        while (ntr > 8 && (c.ncols + ntr - 1) / ntr < sched.min_blocks) ntr >>= 1;
        c.lg_ntr = ilog2(ntr);
        c.pitch = ntr;
        c.blocks = (int)((c.ncols + ntr - 1) / ntr);
        int thr = ((ntr * L / 8 + 63) / 64) * 64;
        c.threads = std::max(64, std::min(256, thr));
        c.lds = (size_t)nbuf * L * ntr * sizeof(C) + 64 * sizeof(int64_t);
        KWH_REQUIRE(c.lds <= 65536, KWH_EUNSUPPORTED, "axis too long for LDS staging");
        return c;
    }

    double spec_bytes() const { return (double)g.NH * sizeof(C); }
    double mapb(const void* p) const { return p ? (double)g.N * sizeof(T) : 0.0; }
    double real_bytes() const { return (double)g.N * sizeof(T); }

    // ------------------------------------------------------------------ pass wrappers
    void strided(int axis, bool inv, std::initializer_list<void*> fields) {
        std::vector<void*> v(fields);
        strided_v(axis, inv, v);
    }
    // columns per tile for the register-resident kernels: the widest tile whose LDS fits (256
    // lanes in float32, 128 in float64), or the autotuner's choice (half of that)
    int reg_ntr(const RegOps<T>* o) const { return o->ntr_max >> sched.ntr_shift; }
    static size_t reg_col_lds(const RegOps<T>* o, int ntr) {
        const int L = o->J * o->E;
        return ((size_t)L * ntr + o->J + L) * sizeof(C);   // slab + pad + twiddle table
    }
    // dual_field >= 0 (register inverse passes only): that field F also yields dual_out =
    // IFFT(d2[column] F) and itself becomes IFFT(d1[n] F) (StridedArgs::dual)
    void strided_v(int axis, bool inv, const std::vector<void*>& fields, int dual_field = -1, void* dual_out = nullptr,
                   const void* d1 = nullptr, const void* d2 = nullptr) {
        SCfg c = scfg(axis, 2);
        if (c.L <= 1) return;
        StridedArgs a{};
        int nf = 0;
        for (void* f : fields) a.fields[nf++] = f;
        a.nfields = nf;
        a.axis = axis;
        a.L = c.L;
        a.stride = c.stride;
        a.inner = c.inner;
        a.ncols = c.ncols;
        double bytes = 2.0 * spec_bytes() * nf;
        if (dual_field >= 0) {
            KWH_REQUIRE(inv && rx[axis] && dual_field < nf, KWH_EUNSUPPORTED, "two-output pass needs a register inverse");
            a.dual = dual_field + 1;
            a.dual_out = dual_out;
            a.d1 = d1;
            a.d2 = d2;
            bytes += spec_bytes();
        }
        a.walk_rev = sched.strided_rev;
        if (const RegOps<T>* o = rx[axis]) {
            const int ntr = reg_ntr(o);
            a.lg_ntr = ilog2(ntr);
            a.pitch = ntr;
            dim3 grid((unsigned)((c.ncols + ntr - 1) / ntr), nf);
            timed(inv ? K_SI : K_SF, bytes, [&] {
                o->strided(em(), grid, dim3(ntr * o->J), reg_col_lds(o, ntr), a, tab.tw[axis], inv, ntr);
            });
            return;
        }
        a.lg_ntr = c.lg_ntr;
        a.pitch = c.pitch;
        dim3 grid(c.blocks, nf);
        timed(inv ? K_SI : K_SF, bytes, [&] {
            if (inv)
                emit<StridedB<T, true>>(grid, dim3(c.threads), c.lds, a, tab.plan[axis], tab.tw[axis]);
            else
                emit<StridedB<T, false>>(grid, dim3(c.threads), c.lds, a, tab.plan[axis], tab.tw[axis]);
        });
    }

    // slab_n > 0: input 0 is zero outside the middle-axis slabs [slab_lo, slab_lo + slab_n); the
    // register pass then forms its forward transform from those slabs (up to kSlabDirect of them)
    static constexpr int kSlabDirect = 4;
    void middle(int kid, std::vector<const void*> ins, std::vector<MidOp> ops, int slab_lo_ = 0, int slab_n_ = 0) {
        SCfg c = scfg(mid_axis, 3);
        MidArgs a{};
        a.nin = (int)ins.size();
        for (int i = 0; i < a.nin; ++i) a.in[i] = ins[i];
        a.nop = (int)ops.size();
        for (int i = 0; i < a.nop; ++i) a.op[i] = ops[i];
        a.axis = mid_axis;
        a.L = std::max(c.L, 1);
        a.stride = c.stride;
        a.inner = c.inner;
        a.ncols = c.ncols;
        a.scale = 1.0 / (double)g.N;
        const RegOps<T>* o = rx[mid_axis];
        // algorithmic bytes: inputs + outputs + one read of each distinct k-space table
        std::vector<const void*> tabs_used;
        for (int i = 0; i < a.nop; ++i) {
            const MidOp& o = a.op[i];
            if (o.table && std::find(tabs_used.begin(), tabs_used.end(), o.table) == tabs_used.end())
                tabs_used.push_back(o.table);
        }
        const int ntab = (int)tabs_used.size();
        double bytes = spec_bytes() * (a.nin + a.nop) + (double)g.NH * sizeof(T) * ntab;
        if (o && slab_n_ > 0 && slab_n_ <= kSlabDirect) {
            a.slab_lo = slab_lo_;
            a.slab_n = slab_n_;
            bytes -= spec_bytes() * (1.0 - (double)slab_n_ / a.L);
        }
        if (o) {
            for (int f = 0; f < a.nin; ++f) {   // one k-space table per input (loaded once in-kernel)
                const void* t0 = nullptr;
                for (auto& o : ops)
                    if (o.in == f && o.table) {
                        KWH_REQUIRE(!t0 || t0 == o.table, KWH_EUNSUPPORTED, "middle pass: two tables for one input");
                        t0 = o.table;
                    }
            }
            const int ntr = reg_ntr(o);
            a.lg_ntr = ilog2(ntr);
            a.pitch = ntr;
            const unsigned blocks = (unsigned)((c.ncols + ntr - 1) / ntr);
            timed(kid, bytes, [&] {
                o->middle(em(), dim3(blocks), dim3(ntr * o->J), reg_col_lds(o, ntr), a, tab, g, ntr);
            });
            return;
        }
        a.lg_ntr = c.lg_ntr;
        a.pitch = c.pitch;
        timed(kid, bytes, [&] {
            emit<MiddleB<T>>(dim3(c.blocks), dim3(c.threads), c.lds, a, tab, g);
        });
    }

    // Launch configuration of the register-resident z kernels: P row pairs x J lanes.
    struct RegZ { int P, blocks, threads; size_t lds; };
    // Row pairs per workgroup: 256 (FP32) or 128 (FP64) threads, halved until the LDS fits 64 KiB.
    RegZ regz(int nrows_used = -1) const {
        const RegOps<T>* o = rx[2];
        const int L = o->J * o->E;
        RegZ z{};
        z.P = ((sizeof(T) == 8 ? 128 : 256) / o->J) >> sched.p_shift;
        const size_t per_pair = (size_t)(L + o->E + 1) * sizeof(C);
        const size_t twb = (size_t)L * sizeof(C);   // LDS twiddle table
        while (z.P > 1 && z.P * per_pair + twb > 65536) z.P /= 2;
        const int npairs = ((nrows_used < 0 ? g.nrows : nrows_used) + 1) / 2;
        z.blocks = (npairs + z.P - 1) / z.P;
        z.threads = z.P * o->J;
        z.lds = (size_t)z.P * per_pair + twb;
        return z;
    }

    // The velocity, density and EOS row kernels in the half form (RegOps::half_J): one real row per
    // lane group, P rows per workgroup.
    bool zhalf() const { return rx[2] && rx[2]->half_J > 0 && sched.row_half; }
    RegZ regz_half() const {
        const RegOps<T>* o = rx[2];
        const int L = o->half_J * o->half_E;   // complex points per transform (rows of 2L)
        RegZ z{};
        z.P = ((sizeof(T) == 8 ? 128 : 256) / o->half_J) >> sched.p_shift;
        const size_t per_row = (size_t)(L + o->half_E + 1) * sizeof(C);
        const size_t twb = (size_t)L * sizeof(C);
        while (z.P > 1 && z.P * per_row + twb > 65536) z.P /= 2;
        z.blocks = (g.nrows + z.P - 1) / z.P;
        z.threads = z.P * o->half_J;
        z.lds = (size_t)z.P * per_row + twb;
        return z;
    }

    // Full forward transform of a real field into `out` (z, then plain passes): leaves the
    // spectrum ready for a middle pass.
    void forward_to_middle(const T* src, C* out, T scale) {
        if (plane) {
            timed(K_PFR, real_bytes() + spec_bytes(), [&] {
                px->fwd_real(em(), g.m0, src, out, scale, g, tab);
            });
            return;
        }
        if (rx[2]) {
            const RegZ z = regz();
            timed(K_ZFR, real_bytes() + spec_bytes(), [&] {
                rx[2]->z_forward_real(em(), dim3(z.blocks), dim3(z.threads), z.lds, src, out, scale, g, z.P, tab);
            });
            if (plain1) strided(1, false, {out});
            return;
        }
        ZCfg zc = zcfg(2);
        timed(K_ZFR, real_bytes() + spec_bytes(), [&] {
            emit<ZForwardRealB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, src, out, scale, g,
                               zc.zt, tab);
        });
        if (plain1) strided(1, false, {out});
    }
    void inverse_from_middle(C* in, T* dst, T scale, bool accumulate) {
        if (plane) {
            timed(K_PIR, real_bytes() * (accumulate ? 2 : 1) + spec_bytes(), [&] {
                px->inv_real(em(), g.m0, (const C*)in, dst, scale, accumulate ? 1 : 0, g, tab);
            });
            return;
        }
        if (plain1) strided(1, true, {in});
        if (rx[2]) {
            const RegZ z = regz();
            timed(K_ZIR, real_bytes() + spec_bytes(), [&] {
                rx[2]->z_inverse_real(em(), dim3(z.blocks), dim3(z.threads), z.lds, (const C*)in, dst, scale,
                                      accumulate ? 1 : 0, g, z.P, tab);
            });
            return;
        }
        ZCfg zc = zcfg(2);
        timed(K_ZIR, real_bytes() + spec_bytes(), [&] {
            emit<ZInverseRealB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, in, dst, scale,
                               accumulate ? 1 : 0, g, zc.zt, tab);
        });
    }

    template <typename U> void fill(U* ptr, U v, int64_t n) {
        timed(K_FI, (double)n * sizeof(U), [&] {
            emit<FillB<U>>(dim3(grid_blocks(n)), dim3(256), 0, ptr, v, n);
        });
    }

    // ------------------------------------------------------------------ setup
    void setup(const kwh_sim_desc& d, int device) {
        dev = device;
        HIPCHK(hipSetDevice(dev));
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        if (const char* e = std::getenv("KWH_LDS_BUDGET")) sched.lds = (size_t)std::atol(e);
        if (const char* e = std::getenv("KWH_MIN_BLOCKS")) sched.min_blocks = std::max(1, std::atoi(e));
        if (const char* e = std::getenv("KWH_NTR_SHIFT")) sched.ntr_shift = std::atoi(e) != 0 ? 1 : 0;
        if (const char* e = std::getenv("KWH_P_SHIFT")) sched.p_shift = std::max(0, std::min(2, std::atoi(e)));
        if (const char* e = std::getenv("KWH_STRIDED_REV")) sched.strided_rev = std::atoi(e) != 0 ? 1 : 0;
        if (const char* e = std::getenv("KWH_GRAD_DUAL")) sched.grad_dual = std::atoi(e) != 0 ? 1 : 0;
        if (const char* e = std::getenv("KWH_CORR_FUSE")) sched.corr_fuse = std::atoi(e) != 0 ? 1 : 0;
        if (const char* e = std::getenv("KWH_ROW_HALF")) sched.row_half = std::atoi(e) != 0 ? 1 : 0;
        scal_pool.alloc((size_t)64 * kScalarRep * sizeof(T));
        zero_map = mk_scalar(T(0));
        taus = etas = bonas = zero_map;
        dim = d.dim;
        if (axisym) validate_axisymmetric(d);
        for (int a = 0; a < dim; ++a) { nphys[a] = d.n[a]; dphys[a] = d.d[a]; }
        dt = d.dt;
        nt = d.nt;
        g.dim = dim;
        g.off = 3 - dim;
        int m[3] = {1, 1, 1};
        for (int a = 0; a < dim; ++a) m[a + g.off] = nphys[a];
        g.m0 = m[0]; g.m1 = m[1]; g.m2 = m[2];
        g.mzh = g.m2 / 2 + 1;
        g.N = (int64_t)g.m0 * g.m1 * g.m2;
        g.NH = (int64_t)g.m0 * g.m1 * g.mzh;
        g.nrows = g.m0 * g.m1;
        mid_axis = g.m0 > 1 ? 0 : 1;
        plain1 = mid_axis == 0 && g.m1 > 1;
        if (const char* e = std::getenv("KWH_GENERIC_FFT")) use_sq = std::atoi(e) == 0;
        {
            const int mm[3] = {g.m0, g.m1, g.m2};
            for (int ax = 0; ax < 3; ++ax)
                rx[ax] = use_sq && mm[ax] > 1 ? reg_ops<T>(mm[ax]) : nullptr;
        }
        if (const char* e = std::getenv("KWH_PLANE")) plane_wanted = std::atoi(e) != 0;
        if (dim == 3 && !axisym && mid_axis == 0 && use_sq) {
            px = plane_ops<T>(g.m1, g.m2);
            int lds_max = 0;
            HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            if (px && px->lds > (size_t)lds_max) px = nullptr;
            if (px) HIPCHK(px->prepare());
        }
        const int64_t N = g.N;

        // ---- media (KSpaceFirstOrder.swift:573-579)
        auto upload_field = [&](const kwh_field& f, DBuf& buf, SMap<T>& sm, const char* name) {
            if (f.count == 1) {
                sm = mk_scalar(T(f.data[0]));
                return;
            }
            KWH_REQUIRE(f.count == N, KWH_EINVAL, std::string(name) + " must be scalar or grid-shaped");
            DBuf tmp;
            tmp.alloc(N * sizeof(float));
            HIPCHK(hipMemcpyAsync(tmp.p, f.data, N * sizeof(float), hipMemcpyHostToDevice, stream));
            buf.alloc(N * sizeof(T));
            hipLaunchKernelGGL((convert_kernel<float, T>), dim3(grid_blocks(N)), dim3(256), 0, stream, tmp.as<float>(),
                               buf.as<T>(), T(1), N);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
            sm = mk_map(buf.as<T>());
        };
        upload_field(d.sound_speed, c0m, c0s, "soundSpeed");
        upload_field(d.density, rho0m, rho0s, "density");

        build_axis_tables_fft();
        if (d.smooth_c0 && c0s.p) smooth_inplace(c0m.as<T>());
        if (d.smooth_rho0 && rho0s.p) smooth_inplace(rho0m.as<T>());
        smoothed_media = (d.smooth_c0 && c0s.p) || (d.smooth_rho0 && rho0s.p);

        // c_ref = max(c0) (:786-788)
        c_ref = c0s.p ? max_abs(c0m.as<T>(), N) : (double)c0s.s;
        // Validate after smoothing as well.
        KWH_REQUIRE(c_ref > 0 && std::isfinite(c_ref), KWH_EINVAL, "soundSpeed must be positive and finite");

        build_axis_tables_ops(d);
        build_kspace_tables(d);
        if (axisym) build_expanded();
        build_media(d);
        build_sources(d);
        build_sensor(d);
        allocate_state(d);
        plane = plane_wanted && plane_supported();   // spec[0] is zero: no conversion
        HIPCHK(hipStreamSynchronize(stream));
    }

    double max_abs(const T* x, int64_t n) {
        DBuf r;
        r.alloc(sizeof(unsigned long long));
        HIPCHK(hipMemsetAsync(r.p, 0, sizeof(unsigned long long), stream));
        hipLaunchKernelGGL((maxabs_kernel<T>), dim3(std::min(grid_blocks(n), 1024)), dim3(256), 0, stream, x, n,
                           r.as<unsigned long long>());
        HIPCHK(hipGetLastError());
        unsigned long long bits = 0;
        HIPCHK(hipMemcpyAsync(&bits, r.p, sizeof(bits), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        double v;
        std::memcpy(&v, &bits, sizeof(v));
        return v;
    }
    double max_abs64(const double* x, int64_t n) {
        DBuf r;
        r.alloc(sizeof(unsigned long long));
        HIPCHK(hipMemsetAsync(r.p, 0, sizeof(unsigned long long), stream));
        hipLaunchKernelGGL((maxabs_kernel<double>), dim3(std::min(grid_blocks(n), 1024)), dim3(256), 0, stream, x, n,
                           r.as<unsigned long long>());
        HIPCHK(hipGetLastError());
        unsigned long long bits = 0;
        HIPCHK(hipMemcpyAsync(&bits, r.p, sizeof(bits), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        double v;
        std::memcpy(&v, &bits, sizeof(v));
        return v;
    }

    template <typename U>
    void upload(DBuf& b, const std::vector<U>& v) {
        b.alloc(v.size() * sizeof(U));
        if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // Twiddles and plans for T and for the float64 setup transforms.
    void build_axis_tables_fft() {
        const int m[3] = {g.m0, g.m1, g.m2};
        for (int ax = 0; ax < 3; ++ax) {
            const int L = m[ax];
            FftPlan pl = L > 1 ? make_plan(L) : FftPlan{1, 0, {}};
            tab.plan[ax] = pl;
            tab64.plan[ax] = pl;
            std::vector<C> tw(L);
            std::vector<cx<double>> tw64(L);
            for (int i = 0; i < L; ++i) {
                const double ang = -2.0 * M_PI * (double)i / (double)L;
                tw[i] = C{T(std::cos(ang)), T(std::sin(ang))};
                tw64[i] = cx<double>{std::cos(ang), std::sin(ang)};
            }
            upload(tw_buf[ax], tw);
            upload(tw64_buf[ax], tw64);
            tab.tw[ax] = tw_buf[ax].as<C>();
            tab64.tw[ax] = tw64_buf[ax].as<cx<double>>();
            // wavenumbers (KWaveGrid.swift:72-78); dummy axes have k = 0
            std::vector<double> kv(L, 0.0);
            if (L > 1 || ax >= g.off) {
                const int a = ax - g.off;
                if (a >= 0) {
                    auto f = fft_freq(L, dphys[a]);
                    for (int i = 0; i < L; ++i) kv[i] = 2.0 * M_PI * f[i];
                }
            }
            upload(kv64_buf[ax], kv);
        }
    }

    // 1-D operator tables with the Hermitian part at an even-length Nyquist bin.
    void build_axis_tables_ops(const kwh_sim_desc& d) {
        const int m[3] = {g.m0, g.m1, g.m2};
        for (int ax = 0; ax < 3; ++ax) {
            const int L = m[ax];
            const int a = ax - g.off;
            std::vector<C> dp(L), dn(L), co(L);
            std::vector<T> pm(L, T(1)), pms(L, T(1));
            if (a >= 0) {
                auto f = fft_freq(L, dphys[a]);
                for (int i = 0; i < L; ++i) {
                    const double k = 2.0 * M_PI * f[i];
                    const double th = k * dphys[a] / 2.0;
                    // i k e^{+i th} = (-k sin th, k cos th); i k e^{-i th} = (k sin th, k cos th)
                    double dpr = -k * std::sin(th), dpi = k * std::cos(th);
                    double dnr = k * std::sin(th), dni = k * std::cos(th);
                    double cr = std::cos(th), ci = -std::sin(th);
                    if (L % 2 == 0 && i == L / 2) { dpi = 0; dni = 0; ci = 0; }
                    dp[i] = C{T(dpr), T(dpi)};
                    dn[i] = C{T(dnr), T(dni)};
                    co[i] = C{T(cr), T(ci)};
                }
                KWH_REQUIRE(d.pml_size >= 0 && d.pml_size <= L, KWH_EINVAL, "pmlSize must be between 0 and the grid size");
                const bool radial = axisym && ax == 2;
                auto p1 = pml_profile(L, dphys[a], dt, c_ref, d.pml_size, d.pml_alpha, false, radial);
                auto p2 = pml_profile(L, dphys[a], dt, c_ref, d.pml_size, d.pml_alpha, true, radial);
                for (int i = 0; i < L; ++i) { pm[i] = T(p1[i]); pms[i] = T(p2[i]); }
            } else {
                for (int i = 0; i < L; ++i) { dp[i] = C{0, 0}; dn[i] = C{0, 0}; co[i] = C{1, 0}; }
            }
            upload(ddxp_buf[ax], dp);
            upload(ddxn_buf[ax], dn);
            upload(coll_buf[ax], co);
            upload(pml_buf[ax], pm);
            upload(pmlsg_buf[ax], pms);
            tab.ddxp[ax] = ddxp_buf[ax].as<C>();
            tab.ddxn[ax] = ddxn_buf[ax].as<C>();
            tab.coll[ax] = coll_buf[ax].as<C>();
            tab.pml[ax] = pml_buf[ax].as<T>();
            tab.pmlsg[ax] = pmlsg_buf[ax].as<T>();
        }
    }

    KGeo kgeo() const {
        KGeo kg{};
        kg.m0 = g.m0; kg.m1 = g.m1; kg.m2 = g.m2; kg.mzh = g.mzh; kg.NH = g.NH;
        kg.kv0 = kv64_buf[0].as<double>();
        kg.kv1 = kv64_buf[1].as<double>();
        kg.kv2 = kv64_buf[2].as<double>();
        return kg;
    }

    void build_kspace_tables(const kwh_sim_desc& d) {
        const bool need_sk = (d.p_source.index && d.p_source.mode == KWH_SRC_ADDITIVE) ||
                             [&] { for (int a = 0; a < 3; ++a) if (d.u_source[a].index && d.u_source[a].signal && d.u_source[a].mode == KWH_SRC_ADDITIVE) return true; return false; }();
        kappa.alloc(g.NH * sizeof(T));
        if (need_sk) skappa.alloc(g.NH * sizeof(T));
        // absorption flags (makeAbsorption, :855-878)
        absorbing = d.alpha_coeff.data != nullptr && d.alpha_mode != KWH_ALPHA_NONE;
        double y = d.alpha_power;
        if (d.alpha_mode == KWH_ALPHA_STOKES) y = 2.0;
        stokes = absorbing && (d.alpha_mode == KWH_ALPHA_STOKES || std::fabs(y - 2.0) < 1e-10);
        dispersion = absorbing && !stokes && d.alpha_mode != KWH_ALPHA_NO_DISPERSION;
        if (absorbing && !stokes) nab1.alloc(g.NH * sizeof(T));
        if (dispersion) nab2.alloc(g.NH * sizeof(T));
        hipLaunchKernelGGL((kspace_tables_kernel<T>), dim3(grid_blocks(g.NH)), dim3(256), 0, stream, kgeo(),
                           c_ref * dt / 2.0, y - 2.0, y - 1.0, kappa.as<T>(), skappa.as<T>(), nab1.as<T>(), nab2.as<T>());
        HIPCHK(hipGetLastError());
    }

    void build_media(const kwh_sim_desc& d) {
        const int64_t N = g.N;
        // dt / staggered rho0 (:610-614)
        for (int a = 0; a < dim; ++a) {
            if (rho0s.p) {
                dtrm[a].alloc(N * sizeof(T));
                hipLaunchKernelGGL((stagger_kernel<T>), dim3(grid_blocks(N)), dim3(256), 0, stream, rho0m.as<T>(),
                                   dtrm[a].as<T>(), g.m0, g.m1, g.m2, a + g.off, dt);
                HIPCHK(hipGetLastError());
                dtrs[a] = mk_map(dtrm[a].as<T>());
            } else {
                dtrs[a] = mk_scalar(T(dt / (double)rho0s.s));
            }
        }
        if (absorbing) {
            double y = d.alpha_mode == KWH_ALPHA_STOKES ? 2.0 : d.alpha_power;
            const double db = db2neper_unit(y);
            DBuf alpha_dev;
            T as = 0;
            const T* ap = nullptr;
            if (d.alpha_coeff.count == 1) as = T(d.alpha_coeff.data[0]);
            else {
                KWH_REQUIRE(d.alpha_coeff.count == N, KWH_EINVAL, "alphaCoeff must be scalar or grid-shaped");
                alpha_dev.alloc(N * sizeof(T));
                DBuf tmp;
                tmp.alloc(N * sizeof(float));
                HIPCHK(hipMemcpyAsync(tmp.p, d.alpha_coeff.data, N * sizeof(float), hipMemcpyHostToDevice, stream));
                hipLaunchKernelGGL((convert_kernel<float, T>), dim3(grid_blocks(N)), dim3(256), 0, stream,
                                   tmp.as<float>(), alpha_dev.as<T>(), T(1), N);
                HIPCHK(hipStreamSynchronize(stream));
                ap = alpha_dev.as<T>();
            }
            const bool map = c0s.p || ap;
            const double tanfac = std::tan(M_PI * y / 2.0);
            if (c0s.p && !ap && !stokes) {
                // scalar alphaCoeff over a sound-speed map: the EOS evaluates tau and eta per
                // cell from c0 instead of reading two precomputed maps
                const double a = (double)as * db;
                tau_eta_from_c = true;
                tk = -2.0 * a;
                te = y - 1.0;
                ek = 2.0 * a * tanfac;
                ee = y;
                taus = zero_map;
                etas = zero_map;
            } else if (map) {
                taum.alloc(N * sizeof(T));
                if (dispersion) etam.alloc(N * sizeof(T));
                hipLaunchKernelGGL((absorption_kernel<T>), dim3(grid_blocks(N)), dim3(256), 0, stream, c0s.p, c0s.s,
                                   ap, as, N, db, y, stokes ? 1 : 0, tanfac, taum.as<T>(),
                                   etam.as<T>());
                HIPCHK(hipGetLastError());
                HIPCHK(hipStreamSynchronize(stream));
                taus = mk_map(taum.as<T>());
                etas = dispersion ? mk_map(etam.as<T>()) : zero_map;
            } else {
                const double c = c0s.s, a = (double)as * db;
                taus = mk_scalar(T(stokes ? -2.0 * a * c : -2.0 * a * std::pow(c, y - 1.0)));
                etas = mk_scalar(T(2.0 * a * std::pow(c, y) * tanfac));
            }
        }
        // B/A (:902-907): disabled when absent or all zero
        if (d.bona.data) {
            bool any = false;
            for (int64_t i = 0; i < d.bona.count; ++i) any |= d.bona.data[i] != 0.0f;
            if (any) {
                nonlinear = true;
                upload_scalar_or_map(d.bona, bonam, bonas, "bOnA");
            }
        }
    }

    void upload_scalar_or_map(const kwh_field& f, DBuf& buf, SMap<T>& sm, const char* name) {
        if (f.count == 1) { sm = mk_scalar(T(f.data[0])); return; }
        KWH_REQUIRE(f.count == g.N, KWH_EINVAL, std::string(name) + " must be scalar or grid-shaped");
        std::vector<T> h(f.data, f.data + f.count);
        upload(buf, h);
        sm = mk_map(buf.as<T>());
    }

    // Smoothing in float64 (Filters.swift:273-347): window in k-space, then restore max|f|.
    void smooth_inplace(T* f) {
        const int64_t N = g.N;
        DBuf x64, s64, win;
        x64.alloc(N * sizeof(double));
        s64.alloc(g.NH * sizeof(cx<double>));
        win.alloc(g.NH * sizeof(double));
        hipLaunchKernelGGL((convert_kernel<T, double>), dim3(grid_blocks(N)), dim3(256), 0, stream, f, x64.as<double>(), 1.0, N);
        HIPCHK(hipGetLastError());
        const double src_max = max_abs64(x64.as<double>(), N);
        build_window(win.as<double>());
        transform_apply64(x64.as<double>(), s64.as<cx<double>>(), win.as<double>());
        const double out_max = max_abs64(x64.as<double>(), N);
        const double scale = out_max > 0 ? src_max / out_max : 1.0;
        hipLaunchKernelGGL((convert_kernel<double, T>), dim3(grid_blocks(N)), dim3(256), 0, stream, x64.as<double>(), f,
                           T(scale), N);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }

    void build_window(double* out) {
        // rotationProfile (Filters.swift:252-262)
        int Lw = 1;
        for (int a = 0; a < dim; ++a) Lw = std::max(Lw, nphys[a] % 2 ? nphys[a] : nphys[a] + 1);
        std::vector<double> w(Lw, 1.0);
        if (Lw > 1)
            for (int i = 0; i < Lw; ++i) {
                const double x = 2.0 * M_PI * i / (Lw - 1);
                w[i] = 0.42 - 0.5 * std::cos(x) + 0.08 * std::cos(2 * x);
            }
        DBuf wd;
        upload(wd, w);
        WinGeo wg{};
        wg.dim = dim;
        for (int a = 0; a < 3; ++a) wg.n[a] = a < dim ? nphys[a] : 1;
        wg.off = g.off;
        wg.m1 = g.m1;
        wg.mzh = g.mzh;
        wg.NH = g.NH;
        wg.win = wd.as<double>();
        wg.Lw = Lw;
        wg.radius = (Lw - 1) / 2.0;
        wg.place2d = dim == 2 ? 1 : 0;
        hipLaunchKernelGGL((window_kernel<double>), dim3(grid_blocks(g.NH)), dim3(256), 0, stream, wg, out);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }

    // x <- C2R(table * R2C(x)) / N in float64 with the same kernels instantiated for double.
    void transform_apply64(double* x, cx<double>* s, const double* table) {
        using D = double;
        const size_t cs = sizeof(cx<double>);
        // z forward
        {
            const int L = g.m2, npairs = (g.nrows + 1) / 2;
            int ntr = 32;
            while (ntr > 1 && ((size_t)2 * L * (ntr + 1) * cs > sched.lds || ntr / 2 >= npairs)) ntr >>= 1;
            ZTile zt{0, ntr, ilog2(ntr), ntr > 1 ? ntr + 1 : 1};
            int thr = std::max(64, std::min(256, ((ntr * L / 8 + 63) / 64) * 64));
            const size_t lds = (size_t)2 * L * zt.pitch * cs;
            const int blocks = (npairs + ntr - 1) / ntr;
            emit<ZForwardRealB<D>>(dim3(blocks), dim3(thr), lds, x, s, 1.0, g, zt, tab64);
            HIPCHK(hipGetLastError());
            auto strided64 = [&](int axis, bool inv) {
                const int La = axis == 0 ? g.m0 : g.m1;
                if (La <= 1) return;
                StridedArgs a{};
                a.fields[0] = s;
                a.nfields = 1;
                a.axis = axis;
                a.L = La;
                if (axis == 0) { a.stride = (int64_t)g.m1 * g.mzh; a.inner = a.stride; a.ncols = a.stride; }
                else { a.stride = g.mzh; a.inner = g.mzh; a.ncols = (int64_t)g.m0 * g.mzh; }
                int nt_ = 64;
                while (nt_ > 1 && ((size_t)2 * La * nt_ * cs + 64 * 8 > sched.lds || nt_ / 2 >= a.ncols)) nt_ >>= 1;
                a.lg_ntr = ilog2(nt_);
                a.pitch = nt_;
                const int th = std::max(64, std::min(256, ((nt_ * La / 8 + 63) / 64) * 64));
                const size_t l = (size_t)2 * La * nt_ * cs + 64 * 8;
                const int bl = (int)((a.ncols + nt_ - 1) / nt_);
                if (inv) emit<StridedB<D, true>>(dim3(bl, 1), dim3(th), l, a, tab64.plan[axis], tab64.tw[axis]);
                else emit<StridedB<D, false>>(dim3(bl, 1), dim3(th), l, a, tab64.plan[axis], tab64.tw[axis]);
                HIPCHK(hipGetLastError());
            };
            if (plain1) strided64(1, false);
            // middle
            {
                const int ax = mid_axis;
                const int La = ax == 0 ? g.m0 : g.m1;
                MidArgs a{};
                a.in[0] = s;
                a.nin = 1;
                a.op[0] = MidOp{OP_TAB, 0, 0, s, table};
                a.nop = 1;
                a.axis = ax;
                a.L = std::max(La, 1);
                if (ax == 0) { a.stride = (int64_t)g.m1 * g.mzh; a.inner = a.stride; a.ncols = a.stride; }
                else { a.stride = g.mzh; a.inner = g.mzh; a.ncols = (int64_t)g.m0 * g.mzh; }
                int nt_ = 64;
                while (nt_ > 1 && ((size_t)3 * a.L * nt_ * cs + 64 * 8 > sched.lds || nt_ / 2 >= a.ncols)) nt_ >>= 1;
                a.lg_ntr = ilog2(nt_);
                a.pitch = nt_;
                a.scale = 1.0 / (double)g.N;
                const int th = std::max(64, std::min(256, ((nt_ * a.L / 8 + 63) / 64) * 64));
                const size_t l = (size_t)3 * a.L * nt_ * cs + 64 * 8;
                const int bl = (int)((a.ncols + nt_ - 1) / nt_);
                emit<MiddleB<D>>(dim3(bl), dim3(th), l, a, tab64, g);
                HIPCHK(hipGetLastError());
            }
            if (plain1) strided64(1, true);
            emit<ZInverseRealB<D>>(dim3(blocks), dim3(thr), lds, s, x, 1.0, 0, g, zt, tab64);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
        }
    }

    // Sources (KSpaceFirstOrder.swift:616-638 / :405-415).
    void build_source(const kwh_source_field& sf, SourceDev<T>& sd, bool is_p) {
        if (!sf.index || !sf.signal || sf.n_points <= 0) return;
        const int64_t N = g.N;
        KWH_REQUIRE(sf.rows == 1 || sf.rows == sf.n_points, KWH_EINVAL,
                    "source signal must have 1 row or one row per source point");
        KWH_REQUIRE(sf.len >= 0, KWH_EINVAL, "source signal length must be >= 0");
        KWH_REQUIRE(sf.mode >= 0 && sf.mode <= 2, KWH_EINVAL, "invalid source mode");
        std::vector<int64_t> idx(sf.index, sf.index + sf.n_points);
        for (int64_t i = 0; i < sf.n_points; ++i) {
            KWH_REQUIRE(idx[i] >= 0 && idx[i] < N, KWH_EINVAL, "source index outside the grid");
            KWH_REQUIRE(i == 0 || idx[i] > idx[i - 1], KWH_EINVAL, "source indices must be ascending and unique");
        }
        sd.present = true;
        sd.mode = sf.mode;
        sd.rows = sf.rows;
        sd.len = sf.len;
        sd.n = sf.n_points;
        upload(sd.index, idx);
        // c0 at the source points (after smoothing)
        std::vector<double> cpts(sf.n_points, (double)c0s.s);
        if (c0s.p) {
            DBuf out;
            out.alloc(sf.n_points * sizeof(double));
            hipLaunchKernelGGL((gather_values_kernel<T>), dim3(grid_blocks(sf.n_points)), dim3(256), 0, stream, c0s.p,
                               sd.index.template as<int64_t>(), sf.n_points, out.as<double>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(cpts.data(), out.p, sf.n_points * sizeof(double), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        const double nd = dim;
        std::vector<T> base(sf.n_points);
        for (int64_t i = 0; i < sf.n_points; ++i) {
            const double c = cpts[i];
            double b;
            if (is_p) b = sf.mode == KWH_SRC_DIRICHLET ? 1.0 / (nd * c * c) : 2.0 * dt / (nd * c);
            else b = sf.mode == KWH_SRC_DIRICHLET ? 1.0 : 2.0 * c * dt;
            base[i] = T(b);
        }
        for (int a = 0; a < 3; ++a)
            sd.factor[a] = (sf.mode == KWH_SRC_DIRICHLET || a >= dim) ? 1.0 : 1.0 / dphys[a];
        upload(sd.base, base);
        if (sf.delay) {
            // delayed copies of one drive row, expanded into per-point rows on the device
            int maxd = 0;
            for (int64_t i = 0; i < sf.n_points; ++i) maxd = std::max(maxd, (int)sf.delay[i]);
            const int len_eff = sf.len + maxd;
            DBuf sig1, dly, scl;
            upload(sig1, std::vector<float>(sf.signal, sf.signal + sf.len));
            upload(dly, std::vector<int>(sf.delay, sf.delay + sf.n_points));
            if (sf.scale) upload(scl, std::vector<float>(sf.scale, sf.scale + sf.n_points));
            sd.signal.alloc((size_t)sf.n_points * len_eff * sizeof(T));
            hipLaunchKernelGGL((delayed_rows_kernel<T>), dim3(grid_blocks((int64_t)sf.n_points * len_eff)), dim3(256), 0,
                               stream, sig1.as<float>(), sf.len, dly.as<int>(), scl.p ? scl.as<float>() : nullptr,
                               sf.n_points, len_eff, sd.signal.template as<T>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
            sd.rows = (int)sf.n_points;
            sd.len = len_eff;
        } else {
            std::vector<T> sig((size_t)sf.rows * sf.len);
            for (size_t i = 0; i < sig.size(); ++i) sig[i] = T(sf.signal[i]);
            upload(sd.signal, sig);
        }
        if (sf.mode != KWH_SRC_ADDITIVE) {
            std::vector<int> rp(g.nrows + 1, 0);
            for (int64_t i = 0; i < sf.n_points; ++i) rp[idx[i] / g.m2 + 1]++;
            for (int r = 0; r < g.nrows; ++r) rp[r + 1] += rp[r];
            upload(sd.rowptr, rp);
        }
    }

    void build_sources(const kwh_sim_desc& d) {
        build_source(d.p_source, psrc, true);
        for (int a = 0; a < dim; ++a) build_source(d.u_source[a], usrc[a], false);
        for (int a = dim; a < 3; ++a)
            KWH_REQUIRE(d.u_source[a].signal == nullptr || d.u_source[a].index == nullptr, KWH_EINVAL,
                        "velocity source component beyond the grid dimension");
        const bool any_add = psrc.mode == KWH_SRC_ADDITIVE && psrc.present;
        bool any_u_add = false;
        for (int a = 0; a < dim; ++a) any_u_add |= usrc[a].present && usrc[a].mode == KWH_SRC_ADDITIVE;
        if (any_add || any_u_add) {
            scr_real.alloc(g.N * sizeof(T));
            scr_spec.alloc(g.NH * sizeof(C));
        }
        if ((any_add || any_u_add) && !axisym) {
            const int mid_len = mid_axis == 0 ? g.m0 : g.m1;
            const int64_t plane = g.N / mid_len;   // cells per middle-axis index (C-order outermost)
            int64_t lo = mid_len, hi = -1;
            auto span = [&](const kwh_source_field& sf) {
                lo = std::min(lo, sf.index[0] / plane);
                hi = std::max(hi, sf.index[sf.n_points - 1] / plane);   // indices are ascending
            };
            if (any_add) span(d.p_source);
            for (int a = 0; a < dim; ++a)
                if (usrc[a].present && usrc[a].mode == KWH_SRC_ADDITIVE) span(d.u_source[a]);
            if (mid_len > 1 && hi >= lo && 2 * (hi - lo + 1) <= mid_len) {
                slab_lo = (int)lo;
                slab_n = (int)(hi - lo + 1);
                src_spec.alloc(g.NH * sizeof(C));
                fill<C>(src_spec.as<C>(), C{0, 0}, g.NH);
            }
        }
        if (any_add) corr_p.alloc(g.N * sizeof(T));
        for (int a = 0; a < dim; ++a)
            if (usrc[a].present && usrc[a].mode == KWH_SRC_ADDITIVE) corr_u[a].alloc(g.N * sizeof(T));
        // p0 (+ smoothing, :640-645)
        has_p0 = d.p0.data != nullptr;
        if (has_p0) {
            KWH_REQUIRE(d.p0.count == g.N, KWH_EINVAL, "p0 must be grid-shaped");
            for (int64_t i = 0; i < g.N; ++i)
                KWH_REQUIRE(std::isfinite(d.p0.data[i]), KWH_EINVAL, "p0 contains non-finite values");
            std::vector<T> h(d.p0.data, d.p0.data + g.N);
            upload(p0s, h);
            if (d.smooth_p0) smooth_inplace(p0s.as<T>());
            smoothed_p0 = d.smooth_p0 != 0;
        }
        KWH_REQUIRE(has_p0 || psrc.present || usrc[0].present || usrc[1].present || usrc[2].present ||
                        d.p_source.index != nullptr || d.u_source[0].index != nullptr,
                    KWH_EINVAL, "solver requires source.p0 or a time-varying p/u source");
    }

    SparseSrc<T> sparse_of(const SourceDev<T>& s, int a) const {
        SparseSrc<T> r{};
        if (!s.present || s.mode == KWH_SRC_ADDITIVE) return r;
        r.rowptr = s.rowptr.template as<int>();
        r.index = s.index.template as<int64_t>();
        r.signal = s.signal.template as<T>();
        r.base = s.base.template as<T>();
        r.rows = s.rows;
        r.len = s.len;
        r.mode = s.mode == KWH_SRC_DIRICHLET ? 0 : 2;
        r.factor = T(s.factor[a]);
        return r;
    }

    // Sensor (Recording.swift:8-97, Directivity.swift:40-79).
    void build_sensor(const kwh_sim_desc& d) {
        record = d.record;
        rec_p = record & (KWH_REC_P | KWH_REC_P_MAX | KWH_REC_P_MIN | KWH_REC_P_RMS | KWH_REC_I_AVG);
        rec_u = record & (KWH_REC_UX | KWH_REC_UY | KWH_REC_UZ | KWH_REC_U_MAX | KWH_REC_U_RMS | KWH_REC_I_AVG);
        const int64_t N = g.N;
        std::vector<int64_t> idx;
        std::vector<T> wts;
        if (d.sensor_cartesian && d.n_cartesian > 0) {
            cart = true;
            nsens = d.n_cartesian;
            ncorner = 1 << dim;
            idx.resize(nsens * ncorner);
            wts.resize(nsens * ncorner);
            for (int64_t p_ = 0; p_ < nsens; ++p_) {
                int i0[3];
                double fr[3];
                for (int a = 0; a < dim; ++a) {
                    const double gc = (double)d.sensor_cartesian[a * nsens + p_] / dphys[a] + (double)(nphys[a] / 2);
                    KWH_REQUIRE(gc >= 0 && gc <= nphys[a] - 1, KWH_EINVAL, "Cartesian sensor point outside grid bounds");
                    const double f = std::floor(gc);
                    i0[a] = (int)f;
                    fr[a] = gc - f;
                }
                for (int c = 0; c < ncorner; ++c) {
                    int64_t flat = 0;
                    double w = 1.0;
                    for (int a = 0; a < dim; ++a) {
                        const int bit = (c >> a) & 1;
                        const int ai = std::min(i0[a] + bit, nphys[a] - 1);
                        w *= bit ? fr[a] : (1 - fr[a]);
                        flat = flat * nphys[a] + ai;
                    }
                    idx[p_ * ncorner + c] = flat;
                    wts[p_ * ncorner + c] = T(w);
                }
            }
        } else if (d.sensor_index && d.n_sensor > 0) {
            nsens = d.n_sensor;
            idx.assign(d.sensor_index, d.sensor_index + nsens);
            for (int64_t i = 0; i < nsens; ++i) {
                KWH_REQUIRE(idx[i] >= 0 && idx[i] < N, KWH_EINVAL, "sensor index outside the grid");
                KWH_REQUIRE(i == 0 || idx[i] > idx[i - 1], KWH_EINVAL, "sensor indices must be ascending and unique");
            }
        }
        if (nsens == 0) { rec_p = rec_u = false; return; }
        upload(sidx, idx);
        if (cart) upload(sw, wts);
        store_p = rec_p && (record & (KWH_REC_P | KWH_REC_I_AVG));
        store_u = rec_u && (record & (KWH_REC_UX | KWH_REC_UY | KWH_REC_UZ | KWH_REC_I_AVG));
        if (store_p) ser_p.alloc((size_t)nt * nsens * sizeof(T));
        if (rec_p && (record & KWH_REC_P_MAX)) pmax.alloc(nsens * sizeof(T));
        if (rec_p && (record & KWH_REC_P_MIN)) pmin.alloc(nsens * sizeof(T));
        if (rec_p && (record & KWH_REC_P_RMS)) psum2.alloc(nsens * sizeof(double));
        // The reference samples every component when any velocity output is requested
        // (Recording.swift:15-17); only the components an output reads are computed here.
        const bool all_u = record & (KWH_REC_U_MAX | KWH_REC_U_RMS | KWH_REC_I_AVG);
        int slot = dim;
        for (int a = 0; a < dim && rec_u; ++a) {
            need_u[a] = all_u || (record & (KWH_REC_UX << a));
            if (!need_u[a]) continue;
            if (record & ((KWH_REC_UX << a) | KWH_REC_I_AVG)) ser_u[a].alloc((size_t)nt * nsens * sizeof(T));
            if (record & KWH_REC_U_MAX) umax[a].alloc(nsens * sizeof(T));
            if (record & KWH_REC_U_RMS) usum2[a].alloc(nsens * sizeof(double));
            ucol[a].alloc(N * sizeof(T));
            ucol_slot[a] = slot++;
        }
        if (need_u[dim - 1]) {
            std::vector<int> rl;
            rl.reserve(idx.size());
            for (int64_t f : idx) rl.push_back((int)(f / g.m2));
            std::sort(rl.begin(), rl.end());
            rl.erase(std::unique(rl.begin(), rl.end()), rl.end());
            nzrows = (int)rl.size();
            upload(zrows, rl);
        }
        // 2D directivity on binary masks
        if (d.directivity_angle && dim == 2 && !cart && rec_p) {
            std::vector<double> ang(nsens);
            for (int64_t i = 0; i < nsens; ++i) ang[i] = (double)d.directivity_angle[idx[i]];
            std::vector<double> uniq(ang);
            std::sort(uniq.begin(), uniq.end());
            uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
            const double size = d.directivity_size > 0 ? d.directivity_size : 10.0 * std::max(dphys[0], dphys[1]);
            ndir = (int)uniq.size();
            DBuf kx, ky;
            {
                auto fx = fft_freq(nphys[0], dphys[0]);
                auto fy = fft_freq(nphys[1], dphys[1]);
                for (auto& v : fx) v *= 2 * M_PI;
                for (auto& v : fy) v *= 2 * M_PI;
                upload(kx, fx);
                upload(ky, fy);
            }
            for (int k = 0; k < ndir; ++k) {
                auto w = std::make_unique<DBuf>();
                w->alloc(g.NH * sizeof(T));
                DirGeo dg{};
                dg.nx = nphys[0]; dg.ny = nphys[1]; dg.mzh = g.mzh; dg.NH = g.NH;
                dg.kx = kx.as<double>(); dg.ky = ky.as<double>();
                dg.ct = std::cos(uniq[k]); dg.st = std::sin(uniq[k]);
                dg.half_size = size / 2; dg.gradient = d.directivity_pattern == KWH_DIR_GRADIENT;
                hipLaunchKernelGGL((directivity_kernel<T>), dim3(grid_blocks(g.NH)), dim3(256), 0, stream, dg, w->as<T>());
                HIPCHK(hipGetLastError());
                std::vector<int64_t> gi, gp;
                for (int64_t i = 0; i < nsens; ++i)
                    if (ang[i] == uniq[k]) { gi.push_back(idx[i]); gp.push_back(i); }
                auto bi = std::make_unique<DBuf>();
                auto bp = std::make_unique<DBuf>();
                upload(*bi, gi);
                upload(*bp, gp);
                dir_count.push_back((int64_t)gi.size());
                dir_w.push_back(std::move(w));
                dir_idx.push_back(std::move(bi));
                dir_pos.push_back(std::move(bp));
            }
            HIPCHK(hipStreamSynchronize(stream));
            dir_field.alloc(N * sizeof(T));
            if (!scr_spec.p) scr_spec.alloc(g.NH * sizeof(C));
        }
    }

    void allocate_state(const kwh_sim_desc&) {
        const int64_t N = g.N;
        p.alloc(N * sizeof(T));
        for (int a = 0; a < dim; ++a) {
            u[a].alloc(N * sizeof(T));
            rho[a].alloc(N * sizeof(T));
        }
        if (absorbing && !stokes) rhoT.alloc(N * sizeof(T));
        int nspec = dim;
        for (int a = 0; a < dim; ++a) if (need_u[a]) ++nspec;
        for (int k = 0; k < std::max(nspec, 2); ++k) spec[k].alloc(g.NH * sizeof(C));
        ctr.alloc(2 * sizeof(int));
        reset_state();
    }

    void reset_state() {
        const int64_t N = g.N;
        fill<T>(p.as<T>(), T(0), N);
        for (int a = 0; a < dim; ++a) {
            fill<T>(u[a].as<T>(), T(0), N);
            fill<T>(rho[a].as<T>(), T(0), N);
        }
        fill<C>(spec[0].as<C>(), C{0, 0}, g.NH);
        P_valid = true;
        steps = 0;
        rec_steps = 0;
        std::fill(ckpt_have.begin(), ckpt_have.end(), 0);
    }

    // ------------------------------------------------------------------ the time step
    // Geometry restricted to n slabs of the middle axis (the same rows and columns as the full grid
    // inside them): the contiguous- and strided-axis passes of those slabs, on pointers offset to
    // the first slab.
    struct SlabScope {
        SimT* s;
        Geo saved;
        SlabScope(SimT* owner, int n) : s(owner), saved(owner->g) {
            Geo& g = s->g;
            if (s->mid_axis == 0) g.m0 = n;
            else g.m1 = n;
            g.nrows = g.m0 * g.m1;
            g.N = (int64_t)g.nrows * g.m2;
            g.NH = (int64_t)g.nrows * g.mzh;
        }
        ~SlabScope() { s->g = saved; }
        SlabScope(const SlabScope&) = delete;
        SlabScope& operator=(const SlabScope&) = delete;
    };
    // dst == nullptr: stop after the x pass and leave the correction's y-z spectrum in scr_spec
    // (plane mode: the velocity plane kernel inverts it itself)
    void compute_correction(SourceDev<T>& s, int t, T* dst, double factor) {
        // Re ifft(source_kappa * fft(scatter(v))) (SourceOp.apply :167-171)
        const int64_t plane = slab_n ? g.N / (mid_axis == 0 ? g.m0 : g.m1) : 0;
        T* sr = scr_real.as<T>() + (int64_t)slab_lo * plane;
        fill<T>(sr, T(0), slab_n ? (int64_t)slab_n * plane : g.N);
        timed(K_SC, 0, [&] {
            emit<ScatterB<T>>(dim3(grid_blocks(s.n)), dim3(256), 0, scr_real.as<T>(),
                               s.index.template as<int64_t>(), s.signal.template as<T>(), s.base.template as<T>(), s.rows, s.len, (int)s.n,
                               (const int*)ctr.p, T(factor));
        });
        if (slab_n) {
            // only the source slabs are transformed along the contiguous and strided axes; the
            // middle pass reads all of src_spec (zero elsewhere) and writes the dense scr_spec
            C* ss = src_spec.as<C>() + (int64_t)slab_lo * (plane / g.m2) * g.mzh;
            {
                SlabScope sc(this, slab_n);
                forward_to_middle(sr, ss, T(1));
            }
            middle(K_MT, {src_spec.p}, {MidOp{OP_TAB, 0, 0, scr_spec.p, skappa.p}}, slab_lo, slab_n);
        } else {
            forward_to_middle(scr_real.as<T>(), scr_spec.as<C>(), T(1));
            middle(K_MT, {scr_spec.p}, {MidOp{OP_TAB, 0, 0, scr_spec.p, skappa.p}});
        }
        if (dst) inverse_from_middle(scr_spec.as<C>(), dst, T(1), false);
    }

    void init_step() {
        // t = 0 override with p0 (:724-733)
        C* P = spec[0].as<C>();
        forward_to_middle(p0s.as<T>(), P, T(1));
        std::vector<MidOp> ops;
        for (int a = 0; a < dim; ++a)
            ops.push_back(MidOp{OP_DDXP_KAPPA, a + g.off, 0, spec[a].p, kappa.p});
        middle(K_MG, {P}, ops);
        if (plain1) {
            if (dim == 3) strided(1, true, {spec[0].p, spec[1].p, spec[2].p});
        }
        InitArgs<T> ia{};
        for (int a = 0; a < dim; ++a) {
            ia.G[a] = spec[a].as<C>();
            ia.u[a] = u[a].as<T>();
            ia.rho[a] = rho[a].as<T>();
            ia.dtr[a] = dtrs[a];
        }
        ia.c0 = c0s;
        ia.p0 = p0s.as<T>();
        ia.p = p.as<T>();
        ia.outP = spec[0].as<C>();
        ZCfg zc = zcfg(2);
        timed(K_ZI, real_bytes() * (1 + 3 * dim) + spec_bytes() * (dim + 1), [&] {
            emit<ZInitB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, ia, g, zc.zt, tab);
        });
        // collocated velocity of the initial state for the t = 0 sample (:735-745)
        for (int a = 0; a < dim; ++a) {
            if (!need_u[a]) continue;
            C* s = spec[ucol_slot[a]].as<C>();
            forward_to_middle(u[a].as<T>(), s, T(1));
            middle(K_MT, {s}, {MidOp{OP_COLL, a + g.off, 0, s, nullptr}});
            inverse_from_middle(s, ucol[a].as<T>(), T(1), false);
        }
        if (plane) strided(1, false, {spec[0].p});   // R2C_z(p) -> R2C_yz(p)
        P_valid = true;
    }

    // 3-D split passes with register kernels on the x and y axes: the gradient's y and z
    // derivatives are applied in the y inverse (strided_v two-output form)
    bool grad_dual() const { return dim == 3 && plain1 && mid_axis == 0 && rx[0] && rx[1] && !axisym && sched.grad_dual; }

    void main_step(int t) {
        // k-space-corrected additive sources for this step
        const T* cu[3] = {nullptr, nullptr, nullptr};
        auto u_active = [&](int a) { return usrc[a].present && usrc[a].mode == KWH_SRC_ADDITIVE && t < usrc[a].len; };
        int nu = 0;
        for (int a = 0; a < dim; ++a) nu += u_active(a) ? 1 : 0;
        const bool p_active = psrc.present && psrc.mode == KWH_SRC_ADDITIVE && t < psrc.len;
        // plane mode with one corrected velocity source and no pressure correction: the velocity
        // plane kernel inverts the correction's spectrum itself (no dense real field)
        const bool fuse = plane && px->corr_fuse && sched.corr_fuse && nu == 1 && !p_active;
        int corr_axis = -1;
        for (int a = 0; a < dim; ++a)
            if (u_active(a)) {
                if (fuse) {
                    compute_correction(usrc[a], t, nullptr, usrc[a].factor[a]);
                    corr_axis = a;
                } else {
                    compute_correction(usrc[a], t, corr_u[a].as<T>(), usrc[a].factor[a]);
                    cu[a] = corr_u[a].as<T>();
                }
            }
        const T* cp = nullptr;
        if (psrc.present && psrc.mode == KWH_SRC_ADDITIVE && t < psrc.len) {
            // base scale 2 dt / (dim c); the 1/d_a factor is applied per axis in z_density
            compute_correction(psrc, t, corr_p.as<T>(), 1.0);
            cp = corr_p.as<T>();
        }
        if (plane) {
            main_step_plane(cu, cp, corr_axis);
            return;
        }
        C* P = spec[0].as<C>();
        // velocity: grad p (velStep :669-678)
        if (plain1) strided(1, false, {P});
        if (grad_dual()) {
            // the x pass writes Gx = IFFT_x(d_x kappa X) and Q = IFFT_x(kappa X); the y inverse forms
            // G_y = IFFT_y(d_y Q) and G_z = IFFT_y(d_z Q) from one read of Q (d_y, d_z do not vary
            // along x): one spectrum written and read less than three x-pass outputs
            middle(K_MG, {P}, {MidOp{OP_DDXP_KAPPA, 0, 0, spec[0].p, kappa.p}, MidOp{OP_TAB, 0, 0, spec[1].p, kappa.p}});
            strided_v(1, true, {spec[0].p, spec[1].p}, 1, spec[2].p, tab.ddxp[1], tab.ddxp[2]);
        } else {
            std::vector<MidOp> ops;
            for (int a = 0; a < dim; ++a) ops.push_back(MidOp{OP_DDXP_KAPPA, a + g.off, 0, spec[a].p, kappa.p});
            middle(K_MG, {P}, ops);
            if (plain1) strided(1, true, {spec[0].p, spec[1].p, spec[2].p});
        }
        {
            VelArgs<T> va{};
            for (int a = 0; a < dim; ++a) {
                va.G[a] = spec[a].as<C>();
                va.u[a] = u[a].as<T>();
                va.dtr[a] = dtrs[a];
                va.corr[a] = cu[a] ? mk_map(cu[a]) : zero_map;
                va.sp[a] = sparse_of(usrc[a], a);
            }
            va.ctr = (const int*)ctr.p;
            double bytes = 0;
            for (int a = 0; a < dim; ++a)
                bytes += 2 * spec_bytes() + 2 * real_bytes() + mapb(dtrs[a].p) + (cu[a] ? real_bytes() : 0);
            bool anycorr = false;
            for (int a = 0; a < dim; ++a) anycorr |= cu[a] != nullptr;
            if (zhalf()) {
                const RegZ z = regz_half();
                timed(K_ZV, bytes, [&] {
                    rx[2]->z_velocity_h(em(), dim3(z.blocks), dim3(z.threads), z.lds, va, g, z.P, tab, anycorr);
                });
            } else if (rx[2]) {
                const RegZ z = regz();
                timed(K_ZV, bytes, [&] {
                    rx[2]->z_velocity(em(), dim3(z.blocks), dim3(z.threads), z.lds, va, g, z.P, tab, anycorr);
                });
            } else {
                ZCfg zc = zcfg(2);
                timed(K_ZV, bytes, [&] {
                    emit<ZVelocityB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, va, g, zc.zt, tab);
                });
            }
        }
        // collocated contiguous-axis velocity on the sensor rows only, from R2C_z(u) (:735-745)
        const bool zcol = rx[2] && need_u[dim - 1];
        if (zcol && nzrows > 0) {
            const RegZ z = regz(nzrows);
            const C* s = spec[dim - 1].as<C>();
            T* dst = ucol[dim - 1].as<T>();
            timed(K_ZCR, (double)nzrows * g.m2 * (sizeof(T) + sizeof(C) / 2), [&] {
                rx[2]->z_colloc_rows(em(), dim3(z.blocks), dim3(z.threads), z.lds, s, dst, (const int*)zrows.p,
                                     nzrows, g, z.P, tab);
            });
        }
        auto full_colloc = [&](int a) { return need_u[a] && !(zcol && a == dim - 1); };
        // divergence (rhoStep :680-692)
        if (plain1) strided(1, false, {spec[0].p, spec[1].p, spec[2].p});
        {
            std::vector<const void*> ins;
            std::vector<MidOp> ops;
            for (int a = 0; a < dim; ++a) {
                ins.push_back(spec[a].p);
                ops.push_back(MidOp{OP_DDXN_KAPPA, a + g.off, a, spec[a].p, kappa.p});
                if (full_colloc(a)) ops.push_back(MidOp{OP_COLL, a + g.off, a, spec[ucol_slot[a]].p, nullptr});
            }
            middle(K_MD, ins, ops);
        }
        if (plain1) {
            strided(1, true, {spec[0].p, spec[1].p, spec[2].p});
            for (int a = 0; a < dim; ++a)
                if (full_colloc(a)) strided(1, true, {spec[ucol_slot[a]].p});
        }
        for (int a = 0; a < dim; ++a)
            if (full_colloc(a)) {
                C* s = spec[ucol_slot[a]].as<C>();
                T* dst = ucol[a].as<T>();
                if (rx[2]) {
                    const RegZ z = regz();
                    timed(K_ZIR, real_bytes() + spec_bytes(), [&] {
                        rx[2]->z_inverse_real(em(), dim3(z.blocks), dim3(z.threads), z.lds, (const C*)s, dst, T(1), 0,
                                              g, z.P, tab);
                    });
                    continue;
                }
                ZCfg zc = zcfg(2);
                timed(K_ZIR, real_bytes() + spec_bytes(), [&] {
                    emit<ZInverseRealB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, s, dst, T(1),
                                       0, g, zc.zt, tab);
                });
            }
        const bool fft_abs = absorbing && !stokes;
        {
            RhoArgs<T> ra{};
            for (int a = 0; a < dim; ++a) {
                ra.D[a] = spec[a].as<C>();
                ra.rho[a] = rho[a].as<T>();
                ra.corr_factor[a] = T(psrc.factor[a]);
                ra.sp[a] = sparse_of(psrc, a);
            }
            ra.rho0 = rho0s;
            ra.c0 = c0s;
            ra.bona = bonas;
            ra.tau = taus;
            ra.corr = cp ? mk_map(cp) : zero_map;
            ra.p = p.as<T>();
            ra.outP = spec[0].as<C>();
            ra.rhoT = rhoT.as<T>();
            ra.F1 = spec[0].as<C>();
            ra.F2 = spec[1].as<C>();
            ra.dt = T(dt);
            ra.mode = fft_abs ? 1 : 0;
            ra.nonlinear = nonlinear;
            ra.stokes = stokes;
            ra.dispersion = dispersion;
            ra.ctr = (const int*)ctr.p;
            const int nbuf = 3 + (fft_abs || stokes ? 1 : 0) + (nonlinear ? 1 : 0);
            double bytes = dim * (spec_bytes() + 2 * real_bytes()) + mapb(rho0s.p) + (cp ? real_bytes() : 0);
            if (nonlinear) bytes += mapb(bonas.p);
            if (fft_abs) bytes += real_bytes() + spec_bytes() * (dispersion ? 2 : 1);
            else bytes += real_bytes() + mapb(c0s.p) + spec_bytes() + (stokes ? mapb(taus.p) : 0);
            if (rx[2]) {
                const bool half = zhalf();
                const RegZ z = half ? regz_half() : regz();
                const int mode = fft_abs ? 1 : (stokes ? 2 : 0);
                // long rows: rhoT and F2 in a pass of their own (reg_z_rhot), after the density kernel
                const bool split = mode == 1 && rx[2]->rhot_split && !half;
                const double rt_bytes = real_bytes() + (dispersion ? spec_bytes() : 0);
                timed(K_ZD, split ? bytes - rt_bytes : bytes, [&] {
                    if (half)
                        rx[2]->z_density_h(em(), dim3(z.blocks), dim3(z.threads), z.lds, ra, g, z.P, tab, nonlinear,
                                           mode, cp != nullptr);
                    else
                        rx[2]->z_density(em(), dim3(z.blocks), dim3(z.threads), z.lds, ra, g, z.P, tab, nonlinear,
                                         mode, cp != nullptr);
                });
                if (split)
                    timed(K_ZRT, dim * real_bytes() + rt_bytes, [&] {
                        rx[2]->z_rhot(em(), dim3(z.blocks), dim3(z.threads), z.lds, ra, g, z.P, tab);
                    });
            } else {
                ZCfg zc = zcfg(nbuf);
                timed(K_ZD, bytes, [&] {
                    emit<ZDensityB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, ra, g, zc.zt, tab);
                });
            }
        }
        if (fft_abs) {
            if (plain1) {
                if (dispersion) strided(1, false, {spec[0].p, spec[1].p});
                else strided(1, false, {spec[0].p});
            }
            std::vector<const void*> ins = {spec[0].p};
            std::vector<MidOp> ops = {MidOp{OP_TAB, 0, 0, spec[0].p, nab1.p}};
            if (dispersion) {
                ins.push_back(spec[1].p);
                ops.push_back(MidOp{OP_TAB, 0, 1, spec[1].p, nab2.p});
            }
            middle(K_MA, ins, ops);
            if (plain1) {
                if (dispersion) strided(1, true, {spec[0].p, spec[1].p});
                else strided(1, true, {spec[0].p});
            }
            EosArgs<T> ea{};
            ea.A1 = spec[0].as<C>();
            ea.A2 = dispersion ? spec[1].as<C>() : nullptr;
            ea.rhoT = rhoT.as<T>();
            ea.c0 = c0s;
            ea.rho0 = rho0s;
            ea.tau = taus;
            ea.eta = etas;
            ea.bona = bonas;
            ea.nonlinear = nonlinear;
            ea.tau_eta_from_c = tau_eta_from_c ? 1 : 0;
            ea.tk = T(tk);
            ea.te = T(te);
            ea.ek = T(ek);
            ea.ee = T(ee);
            ea.p = p.as<T>();
            ea.outP = spec[0].as<C>();
            double bytes = spec_bytes() * (dispersion ? 3 : 2) + 2 * real_bytes() + mapb(c0s.p) + mapb(taus.p) +
                           (dispersion ? mapb(etas.p) : 0) + (nonlinear ? mapb(bonas.p) + mapb(rho0s.p) : 0);
            if (zhalf()) {
                const RegZ z = regz_half();
                timed(K_ZE, bytes, [&] {
                    rx[2]->z_eos_h(em(), dim3(z.blocks), dim3(z.threads), z.lds, ea, g, z.P, tab);
                });
            } else if (rx[2]) {
                const RegZ z = regz();
                timed(K_ZE, bytes, [&] {
                    rx[2]->z_eos(em(), dim3(z.blocks), dim3(z.threads), z.lds, ea, g, z.P, tab);
                });
            } else {
                ZCfg zc = zcfg(3);
                timed(K_ZE, bytes, [&] {
                    emit<ZEosB<T>>(dim3(zc.blocks), dim3(zc.threads), zc.lds, ea, g, zc.zt, tab);
                });
            }
        }
        P_valid = true;
    }

    // The step in plane mode: x pass of P (already R2C_yz(p)), plane velocity kernel (y and z
    // inverse, update, y and z forward per axis), x pass of the divergence, plane density kernel
    // (inverse, update, EOS, R2C_yz(p)). Same operations and order as main_step's split passes.
    void main_step_plane(const T* const cu[3], const T* cp, int corr_axis) {
        C* P = spec[0].as<C>();
        // two-output gradient: the x pass writes Gx (spec 0) and Q = IFFT_x(kappa P) (spec 1); the
        // plane velocity kernel applies d_y / d_z to Q on load and stores R2C_yz(u_y) to spec 2 and
        // R2C_yz(u_z) to spec 1 (after its own read of Q), so axis a's spectra live in sb[a]
        const bool gd = sched.grad_dual && dim == 3;
        const int sb[3] = {0, gd ? 2 : 1, gd ? 1 : 2};
        if (gd) {
            middle(K_MG, {P}, {MidOp{OP_DDXP_KAPPA, 0, 0, spec[0].p, kappa.p}, MidOp{OP_TAB, 0, 0, spec[1].p, kappa.p}});
        } else {
            std::vector<MidOp> ops;
            for (int a = 0; a < dim; ++a) ops.push_back(MidOp{OP_DDXP_KAPPA, a + g.off, 0, spec[a].p, kappa.p});
            middle(K_MG, {P}, ops);
        }
        {
            VelArgs<T> va{};
            bool anycorr = false;
            double bytes = 0;
            va.gdual = gd ? 1 : 0;
            if (corr_axis >= 0) {
                va.corr_spec = scr_spec.as<C>();
                va.corr_axis = corr_axis;
                anycorr = true;
                bytes += spec_bytes();
            }
            for (int a = 0; a < dim; ++a) {
                va.G[a] = spec[sb[a]].as<C>();
                va.Gin[a] = spec[gd && a > 0 ? 1 : a].as<C>();
                va.u[a] = u[a].as<T>();
                va.dtr[a] = dtrs[a];
                va.corr[a] = cu[a] ? mk_map(cu[a]) : zero_map;
                va.sp[a] = sparse_of(usrc[a], a);
                anycorr |= cu[a] != nullptr;
                bytes += 2 * spec_bytes() + 2 * real_bytes() + mapb(dtrs[a].p) + (cu[a] ? real_bytes() : 0);
            }
            va.ctr = (const int*)ctr.p;
            timed(K_PV, bytes, [&] { px->velocity(em(), g.m0, va, g, tab, anycorr); });
        }
        {
            std::vector<const void*> ins;
            std::vector<MidOp> ops;
            for (int a = 0; a < dim; ++a) {
                ins.push_back(spec[sb[a]].p);
                ops.push_back(MidOp{OP_DDXN_KAPPA, a + g.off, a, spec[sb[a]].p, kappa.p});
            }
            middle(K_MD, ins, ops);
        }
        {
            RhoArgs<T> ra{};
            for (int a = 0; a < dim; ++a) {
                ra.D[a] = spec[sb[a]].as<C>();
                ra.rho[a] = rho[a].as<T>();
                ra.corr_factor[a] = T(psrc.factor[a]);
                ra.sp[a] = sparse_of(psrc, a);
            }
            const bool fft_abs = absorbing && !stokes;
            const int mode = fft_abs ? 1 : (stokes ? 2 : 0);
            ra.rho0 = rho0s;
            ra.c0 = c0s;
            ra.bona = bonas;
            ra.tau = taus;
            ra.corr = cp ? mk_map(cp) : zero_map;
            ra.p = p.as<T>();
            ra.outP = P;
            ra.rhoT = rhoT.as<T>();
            ra.F1 = spec[0].as<C>();
            ra.F2 = spec[1].as<C>();
            ra.dt = T(dt);
            ra.mode = mode;
            ra.nonlinear = nonlinear;
            ra.stokes = stokes;
            ra.dispersion = dispersion;
            ra.ctr = (const int*)ctr.p;
            double bytes = dim * (spec_bytes() + 2 * real_bytes()) + mapb(rho0s.p) + (cp ? real_bytes() : 0);
            if (fft_abs) bytes += real_bytes() + spec_bytes() * (dispersion ? 2 : 1);
            else bytes += real_bytes() + mapb(c0s.p) + spec_bytes() + (stokes ? mapb(taus.p) : 0);
            if (nonlinear) bytes += mapb(bonas.p) + dim * real_bytes();
            timed(K_PD, bytes, [&] {
                px->density(em(), g.m0, ra, g, tab, nonlinear, mode, cp != nullptr);
            });
            if (fft_abs) {
                // x pass of the absorption inputs (as main_step), then the plane EOS
                std::vector<const void*> ins = {spec[0].p};
                std::vector<MidOp> ops = {MidOp{OP_TAB, 0, 0, spec[0].p, nab1.p}};
                if (dispersion) {
                    ins.push_back(spec[1].p);
                    ops.push_back(MidOp{OP_TAB, 0, 1, spec[1].p, nab2.p});
                }
                middle(K_MA, ins, ops);
                EosArgs<T> ea{};
                ea.A1 = spec[0].as<C>();
                ea.A2 = dispersion ? spec[1].as<C>() : nullptr;
                ea.rhoT = rhoT.as<T>();
                ea.c0 = c0s;
                ea.rho0 = rho0s;
                ea.tau = taus;
                ea.eta = etas;
                ea.bona = bonas;
                ea.nonlinear = nonlinear;
                ea.tau_eta_from_c = tau_eta_from_c ? 1 : 0;
                ea.tk = T(tk);
                ea.te = T(te);
                ea.ek = T(ek);
                ea.ee = T(ee);
                ea.p = p.as<T>();
                ea.outP = P;
                const double eb = spec_bytes() * (dispersion ? 3 : 2) + 2 * real_bytes() + mapb(c0s.p) +
                                  mapb(taus.p) + (dispersion ? mapb(etas.p) : 0) +
                                  (nonlinear ? mapb(bonas.p) + mapb(rho0s.p) + real_bytes() : 0);
                timed(K_PE, eb, [&] { px->eos(em(), g.m0, ea, g, tab); });
            }
        }
        P_valid = true;
    }

    // Plane mode covers 3-D grids whose (m1, m2) plane has plane kernels, every medium, no
    // collocated velocity output (computed from the z-only spectra of the split passes) and no
    // checkpointing.
    bool plane_supported() const {
        if (!px || axisym || ckpt_every > 0) return false;
        for (int a = 0; a < dim; ++a)
            if (need_u[a]) return false;
        return true;
    }
    void set_plane(bool on) override {
        KWH_REQUIRE(!on || plane_supported(), KWH_EUNSUPPORTED,
                    "plane kernels need a 3-D grid with a 32x32, 64x64, 96x96 or 128x128 (y, z) plane, "
                    "no collocated velocity output and no checkpointing");
        plane_wanted = on;
        apply_plane(on);
    }
    // Plane mode follows the request wherever it applies (checkpointing switches it off).
    void refresh_plane() { apply_plane(plane_wanted && plane_supported()); }
    void apply_plane(bool on) {
        if (on == plane) return;
        HIPCHK(hipSetDevice(dev));
        sync();
        // spec[0] between steps: R2C_z(p) in split mode, R2C_yz(p) in plane mode. The transforms
        // are unnormalised, so the way back along y carries the 1/m1 of the inverse.
        if (steps > 0) {
            strided(1, !on, {spec[0].p});
            if (!on) {
                const int64_t n = 2 * g.NH;
                hipLaunchKernelGGL((convert_kernel<T, T>), dim3(grid_blocks(n)), dim3(256), 0, stream, spec[0].as<T>(),
                                   spec[0].as<T>(), T(1.0 / g.m1), n);
                HIPCHK(hipGetLastError());
            }
        }
        plane = on;
        for (auto& sg : step_graphs) {
            if (sg.second.exec) (void)hipGraphExecDestroy(sg.second.exec);
            (void)hipGraphDestroy(sg.second.graph);
        }
        step_graphs.clear();
        step_records.clear();
        sync();
    }
    bool plane_mode() const override { return plane; }
    bool plane_available() const override { return plane_supported(); }

    void record_step(int) {
        if (!(rec_p || rec_u)) return;
        if (rec_p) {
            if (ndir > 0) {
                // directivity: per unique angle Re ifft2(fft2(p) W) sampled at that angle's rows
                forward_to_middle(p.as<T>(), scr_spec.as<C>(), T(1));
                for (int k = 0; k < ndir; ++k) {
                    C* tmp = spec[dim > 1 ? 1 : 0].as<C>();  // free at this point of the step
                    middle(K_MT, {scr_spec.p}, {MidOp{OP_TAB, 0, 0, tmp, dir_w[k]->p}});
                    inverse_from_middle(tmp, dir_field.as<T>(), T(1), false);
                    gather(dir_field.as<T>(), dir_idx[k]->as<int64_t>(), nullptr, 1, dir_count[k], ser_p.as<T>(),
                           pmax.as<T>(), pmin.as<T>(), psum2.as<double>(), dir_pos[k]->as<int64_t>());
                }
            }
        }
        // p (without directivity) and the collocated velocity components in one launch
        GatherSet<T> set{};
        int nf = 0;
        if (rec_p && ndir == 0)
            set.f[nf++] = GatherArgs<T>{p.as<T>(), sidx.as<int64_t>(), sw.as<T>(), ncorner, nsens, ser_p.as<T>(),
                                        pmax.as<T>(), pmin.as<T>(), psum2.as<double>(), (const int*)ctr.p};
        for (int a = 0; a < dim; ++a)
            if (need_u[a])
                set.f[nf++] = GatherArgs<T>{ucol[a].as<T>(), sidx.as<int64_t>(), sw.as<T>(), ncorner, nsens,
                                            ser_u[a].as<T>(), umax[a].as<T>(), nullptr, usum2[a].as<double>(),
                                            (const int*)ctr.p};
        if (nf > 0)
            timed(K_GA, (double)nf * nsens * sizeof(T) * 2, [&] {
                emit<GatherB<T>>(dim3(grid_blocks(nsens), nf), dim3(256), 0, set);
            });
    }

    // Subset gather through output positions (directivity rows).
    void gather(const T* field, const int64_t* idx, const T* w, int nc, int64_t n, T* series, T* vmax, T* vmin,
                double* sum2, const int64_t* outpos) {
        (void)w;
        (void)nc;
        timed(K_GA, 0, [&] {
            emit<GatherPosB<T>>(dim3(grid_blocks(n)), dim3(256), 0, field, idx, outpos, n,
                               nsens, series, vmax, vmin, sum2, (const int*)ctr.p);
        });
    }

    // The launches of step t, ending with the device counter advance. Host-side mirrors
    // (steps, rec_steps, P_valid) are updated by the caller.
    void step_launches(int t) {
        if (axisym) {
            as_step(t);
            record_step(t);
            emit<AdvanceB>(dim3(1), dim3(64), 0, (int*)ctr.p, (rec_p || rec_u) ? 1 : 0);
            HIPCHK(hipGetLastError());
            return;
        }
        if (t == 0 && has_p0) init_step();
        else main_step(t);
        record_step(t);
        const int recorded = (rec_p || rec_u) ? 1 : 0;
        emit<AdvanceB>(dim3(1), dim3(64), 0, (int*)ctr.p, recorded);
        HIPCHK(hipGetLastError());
    }
    void step_done(int t) override {
        if (rec_p || rec_u) ++rec_steps;
        P_valid = true;
        steps = t + 1;
    }

    // What step t launches beyond the fixed step: which k-space-corrected additive sources are
    // active (bits 0-3) and whether it is the t = 0 p0 step (bit 4). Everything else reads t
    // from the device counters, so steps with one key share one graph.
    unsigned step_key(int t) const {
        unsigned key = 0;
        for (int a = 0; a < dim; ++a)
            if (usrc[a].present && usrc[a].mode == KWH_SRC_ADDITIVE && t < usrc[a].len) key |= 1u << a;
        if (psrc.present && psrc.mode == KWH_SRC_ADDITIVE && t < psrc.len) key |= 8u;
        if (t == 0 && has_p0) key |= 16u;
        return key;
    }
    StepGraph& graph_for(int t) {
        const unsigned key = step_key(t);
        for (auto& e : step_graphs)
            if (e.first == key) return e.second;
        StepGraph ng{};
        HIPCHK(hipSetDevice(dev));
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        step_launches(t);
        HIPCHK(hipStreamEndCapture(stream, &ng.graph));
        step_graphs.emplace_back(key, ng);
        return step_graphs.back().second;
    }
    void graph_step(int t) override {
        StepGraph& sg = graph_for(t);
        if (!sg.exec) HIPCHK(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphLaunch(sg.exec, stream));
        step_done(t);
    }
    std::list<StepRecord> step_records;
    const StepRecord& recorded_step(int t) override {
        const unsigned key = step_key(t);
        for (const StepRecord& r : step_records)
            if (r.key == key) return r;
        StepRecord r;
        r.key = key;
        recording = &r.launches;
        try {
            step_launches(t);
        } catch (...) {
            recording = nullptr;
            throw;
        }
        recording = nullptr;
        uint64_t h = 1469598103934665603ull;   // FNV-1a over (kernel, workgroup size, LDS)
        auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
        for (const LaunchRec& l : r.launches) {
            mix((uint64_t)(uintptr_t)l.fn);
            mix(l.block);
            mix(l.lds);
        }
        r.signature = h;
        step_records.push_back(std::move(r));
        return step_records.back();
    }
    int total_steps() const override { return nt; }
    int device() const override { return dev; }
    uint64_t shape_class() const override {
        return ((uint64_t)g.m0 * 65599u + (uint64_t)g.m1) * 65599u * 65599u + (uint64_t)g.m2 * 16u + sizeof(T);
    }
    void set_counters_on(hipStream_t s) override {   // nullptr: the sim's own stream
        HIPCHK(hipSetDevice(dev));
        hipLaunchKernelGGL(set_counters_kernel, dim3(1), dim3(64), 0, s ? s : stream, (int*)ctr.p, steps, rec_steps);
        HIPCHK(hipGetLastError());
    }

    void step(int t) {
        if (use_graph && !timing) {
            StepGraph& sg = graph_for(t);
            if (!sg.exec) HIPCHK(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0));
            HIPCHK(hipGraphLaunch(sg.exec, stream));
        } else {
            step_launches(t);
        }
        step_done(t);
        if (ckpt_every > 0 && t % ckpt_every == 0) save_checkpoint(t / ckpt_every);
    }

    // ------------------------------------------------------------------ axisymmetric solver
    // Preconditions of kspaceFirstOrderAS (KSpaceFirstOrderAS.swift:35-38, :95-101): a 2D grid
    // (x axial, y radial), Stokes absorption only, sources p0 / p / ux / uy; the reference returns
    // the pressure series and pFinal (:299-308), so only pressure fields may be recorded here.
    void validate_axisymmetric(const kwh_sim_desc& d) {
        KWH_REQUIRE(d.dim == 2, KWH_EINVAL, "axisymmetric solver requires a 2D grid (x axial, y radial)");
        if (d.alpha_coeff.data && d.alpha_mode != KWH_ALPHA_NONE) {
            const double y = d.alpha_mode == KWH_ALPHA_STOKES ? 2.0 : d.alpha_power;
            KWH_REQUIRE(std::fabs(y - 2.0) < 1e-10, KWH_EINVAL,
                        "axisymmetric absorption requires the Stokes model (alphaPower = 2)");
        }
        const uint32_t pfields = KWH_REC_P | KWH_REC_P_MAX | KWH_REC_P_MIN | KWH_REC_P_RMS | KWH_REC_P_FINAL;
        KWH_REQUIRE((d.record & ~pfields) == 0, KWH_EUNSUPPORTED,
                    "the axisymmetric solver records the pressure only (p, pMax, pMin, pRms, pFinal)");
        KWH_REQUIRE(!d.directivity_angle, KWH_EUNSUPPORTED, "the axisymmetric solver has no sensor directivity");
        KWH_REQUIRE(!d.u_source[2].index || !d.u_source[2].signal, KWH_EINVAL, "uz source on a 2D grid");
    }

    // The expanded space [1][nx][4 ny]: transform plans and twiddles, the staggered derivative and
    // shift operators of both axes (expanded ky, :63-67), kappa and the source kappa from the
    // expanded |k| with c_ref (:55-57, :165-166), the staggered 1/r table (:85-86) and the
    // expanded work buffers.
    void build_expanded() {
        const int nx = g.m1, ny = g.m2;
        Geo G{};
        G.dim = 2; G.off = 1; G.m0 = 1; G.m1 = nx; G.m2 = 4 * ny;
        G.mzh = G.m2 / 2 + 1;
        G.N = (int64_t)nx * G.m2;
        G.NH = (int64_t)nx * G.mzh;
        G.nrows = nx;
        ex.g = G;
        ex.mid_axis = 1;
        ex.plain1 = false;
        ex.rx[0] = nullptr;
        ex.rx[1] = use_sq && nx > 1 ? reg_ops<T>(nx) : nullptr;
        ex.rx[2] = use_sq ? reg_ops<T>(G.m2) : nullptr;
        const int m[3] = {1, nx, G.m2};
        const double dd[3] = {1.0, dphys[0], dphys[1]};
        for (int ax = 0; ax < 3; ++ax) {
            const int L = m[ax];
            ex.tab.plan[ax] = L > 1 ? make_plan(L) : FftPlan{1, 0, {}};
            std::vector<C> tw(L), dp(L), dn(L), co(L);
            std::vector<double> kv(L, 0.0);
            for (int i = 0; i < L; ++i) {
                const double ang = -2.0 * M_PI * (double)i / (double)L;
                tw[i] = C{T(std::cos(ang)), T(std::sin(ang))};
            }
            if (ax > 0) {
                const auto f = fft_freq(L, dd[ax]);
                for (int i = 0; i < L; ++i) {
                    const double k = 2.0 * M_PI * f[i];
                    kv[i] = k;
                    const double th = k * dd[ax] / 2.0;
                    double dpr = -k * std::sin(th), dpi = k * std::cos(th);
                    double dnr = k * std::sin(th), dni = k * std::cos(th);
                    double cr = std::cos(th), ci = -std::sin(th);
                    if (L % 2 == 0 && i == L / 2) { dpi = 0; dni = 0; ci = 0; }   // Hermitian part
                    dp[i] = C{T(dpr), T(dpi)};
                    dn[i] = C{T(dnr), T(dni)};
                    co[i] = C{T(cr), T(ci)};
                }
            } else {
                dp[0] = C{0, 0};
                dn[0] = C{0, 0};
                co[0] = C{1, 0};
            }
            upload(ex_tw[ax], tw);
            upload(ex_ddxp[ax], dp);
            upload(ex_ddxn[ax], dn);
            upload(ex_coll[ax], co);
            upload(ex_kv[ax], kv);
            ex.tab.tw[ax] = ex_tw[ax].as<C>();
            ex.tab.ddxp[ax] = ex_ddxp[ax].as<C>();
            ex.tab.ddxn[ax] = ex_ddxn[ax].as<C>();
            ex.tab.coll[ax] = ex_coll[ax].as<C>();
            ex.tab.pml[ax] = nullptr;
            ex.tab.pmlsg[ax] = nullptr;
        }
        ex_kappa.alloc(G.NH * sizeof(T));
        ex_skappa.alloc(G.NH * sizeof(T));
        KGeo kg{};
        kg.m0 = 1; kg.m1 = nx; kg.m2 = G.m2; kg.mzh = G.mzh; kg.NH = G.NH;
        kg.kv0 = ex_kv[0].as<double>();
        kg.kv1 = ex_kv[1].as<double>();
        kg.kv2 = ex_kv[2].as<double>();
        hipLaunchKernelGGL((kspace_tables_kernel<T>), dim3(grid_blocks(G.NH)), dim3(256), 0, stream, kg,
                           c_ref * dt / 2.0, 0.0, 0.0, ex_kappa.as<T>(), ex_skappa.as<T>(), (T*)nullptr, (T*)nullptr);
        HIPCHK(hipGetLastError());
        std::vector<T> ir(ny);
        for (int y = 0; y < ny; ++y) ir[y] = T(1.0 / ((y + 0.5) * dphys[1]));
        upload(invr_buf, ir);
        for (int k = 0; k < 3; ++k) {
            ex_real[k].alloc(G.N * sizeof(T));
            ex_spec[k].alloc(G.NH * sizeof(C));
        }
    }

    void as_mirror(const T* w, const T* invr, T* E, int kind) {
        const int nx = g.m1, ny = g.m2;
        hipLaunchKernelGGL((as_mirror_kernel<T>), dim3(grid_blocks((int64_t)nx * 4 * ny)), dim3(256), 0, stream, w, invr,
                           E, nx, ny, kind);
        HIPCHK(hipGetLastError());
    }
    // ex_real[1] = dp/dx, ex_real[2] = dp/dy on the expanded grid from the WSWA mirror of pf
    // (pressureGradients, :216-221).
    void as_gradients(const T* pf) {
        as_mirror(pf, nullptr, ex_real[0].as<T>(), AS_WSWA);
        ExScope sc(this);
        forward_to_middle(ex_real[0].as<T>(), ex_spec[0].as<C>(), T(1));
        middle(K_MG, {ex_spec[0].p}, {MidOp{OP_DDXP_KAPPA, 1, 0, ex_spec[1].p, ex_kappa.p},
                                      MidOp{OP_DDXP_KAPPA, 2, 0, ex_spec[2].p, ex_kappa.p}});
        inverse_from_middle(ex_spec[1].as<C>(), ex_real[1].as<T>(), T(1), false);
        inverse_from_middle(ex_spec[2].as<C>(), ex_real[2].as<T>(), T(1), false);
    }
    void as_scatter(T* d0, T* d1, SourceDev<T>& s, int t, double factor, bool add) {
        if (t >= s.len) return;
        hipLaunchKernelGGL((as_scatter_kernel<T>), dim3(grid_blocks(s.n)), dim3(256), 0, stream, d0, d1,
                           s.index.template as<int64_t>(), s.signal.template as<T>(), s.base.template as<T>(), s.rows,
                           s.len, (int)s.n, t, T(factor), add ? 1 : 0);
        HIPCHK(hipGetLastError());
    }
    // E = Re ifft2(source_kappa fft2(mirror(scatter(v)))) on the expanded grid (correctSource, :170-173).
    void as_corrected(SourceDev<T>& s, int t, int kind, double factor, T* E, C* S) {
        fill<T>(scr_real.as<T>(), T(0), g.N);
        as_scatter(scr_real.as<T>(), nullptr, s, t, factor, false);
        as_mirror(scr_real.as<T>(), nullptr, E, kind);
        ExScope sc(this);
        forward_to_middle(E, S, T(1));
        middle(K_MT, {(const void*)S}, {MidOp{OP_TAB, 0, 0, S, ex_skappa.p}});
        inverse_from_middle(S, E, T(1), false);
    }
    void as_add_crop(T* f0, T* f1, const T* E) {
        const int nx = g.m1, ny = g.m2;
        hipLaunchKernelGGL((as_add_crop_kernel<T>), dim3(grid_blocks(g.N)), dim3(256), 0, stream, f0, f1, E, nx, ny);
        HIPCHK(hipGetLastError());
    }

    // One step of the axisymmetric loop (:236-297).
    void as_step(int t) {
        const int nx = g.m1, ny = g.m2;
        const int blocks = grid_blocks(g.N);
        if (t == 0 && has_p0) {
            // t = 0 leapfrog override with p0 (:289-296); the rest of the t = 0 step is overwritten
            as_gradients(p0s.as<T>());
            hipLaunchKernelGGL((as_init_kernel<T>), dim3(blocks), dim3(256), 0, stream, (const T*)p0s.as<T>(), p.as<T>(),
                               rho[0].as<T>(), rho[1].as<T>(), u[0].as<T>(), u[1].as<T>(), (const T*)ex_real[1].as<T>(),
                               (const T*)ex_real[2].as<T>(), c0s, dtrs[0], dtrs[1], nx, ny);
            HIPCHK(hipGetLastError());
            return;
        }
        // velocity from the pressure gradients (:237-241)
        as_gradients(p.as<T>());
        hipLaunchKernelGGL((as_velocity_kernel<T>), dim3(blocks), dim3(256), 0, stream, u[0].as<T>(), u[1].as<T>(),
                           (const T*)ex_real[1].as<T>(), (const T*)ex_real[2].as<T>(), dtrs[0], dtrs[1],
                           tab.pmlsg[1], tab.pmlsg[2], nx, ny);
        HIPCHK(hipGetLastError());
        // velocity sources: ux mirrors WSWA, uy HAHS (applyVelocitySource, :177-194)
        for (int a = 0; a < 2; ++a) {
            SourceDev<T>& s = usrc[a];
            if (!s.present || t >= s.len) continue;
            if (s.mode == KWH_SRC_ADDITIVE) {
                as_corrected(s, t, a == 0 ? AS_WSWA : AS_HAHS, s.factor[a], ex_real[0].as<T>(), ex_spec[0].as<C>());
                as_add_crop(u[a].as<T>(), nullptr, ex_real[0].as<T>());
            } else {
                as_scatter(u[a].as<T>(), nullptr, s, t, s.factor[a], s.mode == KWH_SRC_ADDITIVE_NO_CORRECTION);
            }
        }
        // divergence: du_x/dx and du_y/dy + u_y / r on the staggered radial grid (:243-251)
        as_mirror(u[0].as<T>(), nullptr, ex_real[0].as<T>(), AS_WSWA);
        as_mirror(u[1].as<T>(), nullptr, ex_real[1].as<T>(), AS_HAHS);
        as_mirror(u[1].as<T>(), invr_buf.as<T>(), ex_real[2].as<T>(), AS_HSHA);
        {
            ExScope sc(this);
            for (int k = 0; k < 3; ++k) forward_to_middle(ex_real[k].as<T>(), ex_spec[k].as<C>(), T(1));
            middle(K_MD, {ex_spec[0].p, ex_spec[1].p, ex_spec[2].p},
                   {MidOp{OP_DDXN_KAPPA, 1, 0, ex_spec[0].p, ex_kappa.p}, MidOp{OP_DDXN_KAPPA, 2, 1, ex_spec[1].p, ex_kappa.p},
                    MidOp{OP_COLL, 2, 2, ex_spec[2].p, ex_kappa.p}});
            inverse_from_middle(ex_spec[0].as<C>(), ex_real[0].as<T>(), T(1), false);
            inverse_from_middle(ex_spec[1].as<C>(), ex_real[1].as<T>(), T(1), false);
            inverse_from_middle(ex_spec[2].as<C>(), ex_real[1].as<T>(), T(1), true);
        }
        // mass conservation (:253-261)
        hipLaunchKernelGGL((as_density_kernel<T>), dim3(blocks), dim3(256), 0, stream, rho[0].as<T>(), rho[1].as<T>(),
                           (const T*)ex_real[0].as<T>(), (const T*)ex_real[1].as<T>(), rho0s, tab.pml[1], tab.pml[2],
                           T(dt), nonlinear ? 1 : 0, nx, ny);
        HIPCHK(hipGetLastError());
        // pressure source into both density splits, scale 1/(2 c^2) or 2 dt/(2 c dx) (:263-279)
        if (psrc.present && t < psrc.len) {
            if (psrc.mode == KWH_SRC_ADDITIVE) {
                as_corrected(psrc, t, AS_WSWA, psrc.factor[0], ex_real[2].as<T>(), ex_spec[2].as<C>());
                as_add_crop(rho[0].as<T>(), rho[1].as<T>(), ex_real[2].as<T>());
            } else {
                as_scatter(rho[0].as<T>(), rho[1].as<T>(), psrc, t, psrc.factor[0],
                           psrc.mode == KWH_SRC_ADDITIVE_NO_CORRECTION);
            }
        }
        // equation of state (:281-287)
        hipLaunchKernelGGL((as_eos_kernel<T>), dim3(blocks), dim3(256), 0, stream, p.as<T>(), (const T*)rho[0].as<T>(),
                           (const T*)rho[1].as<T>(), (const T*)ex_real[0].as<T>(), (const T*)ex_real[1].as<T>(), c0s, rho0s,
                           taus, bonas, stokes ? 1 : 0, nonlinear ? 1 : 0, nx, ny);
        HIPCHK(hipGetLastError());
    }

    // ------------------------------------------------------------------ time reversal
    // Runs the loop with the p source rows reversed in time on the device (the recorded data
    // re-injected as a Dirichlet boundary), then writes max(0, factor * pFinal).
    void time_reversal_result(double factor, void* dst) override {
        KWH_REQUIRE(psrc.present && psrc.mode == KWH_SRC_DIRICHLET, KWH_EINVAL, "time reversal needs its Dirichlet source");
        HIPCHK(hipSetDevice(dev));
        hipLaunchKernelGGL((reverse_rows_kernel<T>), dim3(grid_blocks((int64_t)psrc.rows * std::max(psrc.len / 2, 1))),
                           dim3(256), 0, stream, psrc.signal.template as<T>(), psrc.rows, psrc.len);
        HIPCHK(hipGetLastError());
        launch(0, nt);
        DBuf out;
        out.alloc(g.N * sizeof(T));
        hipLaunchKernelGGL((scaled_positive_kernel<T>), dim3(grid_blocks(g.N)), dim3(256), 0, stream, p.as<T>(),
                           out.as<T>(), T(factor), g.N);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dst, out.p, g.N * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // ------------------------------------------------------------------ autotuner
    // Times one full time step (row, strided and middle passes together) under each launch
    // schedule candidate on this sim's own buffers and keeps the fastest; the choice is cached
    // per grid shape, dimension and precision for later sims in the process. State is reset.
    std::string autotune() override {
        KWH_REQUIRE(steps == 0, KWH_ESTATE, "autotune before the first step");
        KWH_REQUIRE(!axisym, KWH_EUNSUPPORTED, "the autotuner covers kspaceFirstOrder");
        HIPCHK(hipSetDevice(dev));
        sync();
        char keyb[96];
        std::snprintf(keyb, sizeof(keyb), "%d:%d:%d:%d:%zu:%d%d%d", dim, g.m0, g.m1, g.m2, sizeof(T), rx[0] ? 1 : 0,
                      rx[1] ? 1 : 0, rx[2] ? 1 : 0);
        const std::string key = device_tag(dev) + " " + keyb;
        const std::string path = tune_cache_path();
        std::string report;
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            tune_read(path);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                sched.ntr_shift = it->second.ntr_shift;
                sched.p_shift = it->second.p_shift;
                sched.lds = it->second.lds;
                sched.min_blocks = it->second.min_blocks;
                char line[200];
                std::snprintf(line, sizeof(line), "cached schedule for %s: ntr_shift=%d p_shift=%d lds=%zu min_blocks=%d\n",
                              key.c_str(), sched.ntr_shift, sched.p_shift, sched.lds, sched.min_blocks);
                return line;
            }
        }
        const bool reg_cols = sizeof(T) == 4 && (rx[mid_axis] || (plain1 && rx[1]));
        bool generic = false;
        const int m[3] = {g.m0, g.m1, g.m2};
        for (int ax = 0; ax < 3; ++ax) generic |= m[ax] > 1 && !rx[ax];
        std::vector<Schedule> cands;
        for (int ns = 0; ns <= (reg_cols ? 1 : 0); ++ns)
            for (int ps = 0; ps <= (rx[2] ? 1 : 0); ++ps)
                for (int lb = 0; lb <= (generic ? 2 : 0); ++lb) {
                    Schedule c;
                    c.ntr_shift = ns;
                    c.p_shift = ps;
                    c.lds = lb == 1 ? 32 * 1024 : 64 * 1024;
                    c.min_blocks = lb == 2 ? 4096 : 1024;
                    cands.push_back(c);
                }
        for (auto& sg : step_graphs) {
            if (sg.second.exec) (void)hipGraphExecDestroy(sg.second.exec);
            (void)hipGraphDestroy(sg.second.graph);
        }
        step_graphs.clear();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const bool tm = timing;
        timing = false;
        float best = 1e30f;
        Schedule chosen = sched;
        for (const Schedule& c : cands) {
            sched = c;
            reset_state();
            main_step(1);
            HIPCHK(hipEventRecord(e0, stream));
            for (int i = 0; i < 3; ++i) main_step(1);
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            ms /= 3;
            char line[160];
            std::snprintf(line, sizeof(line), "ntr_shift=%d p_shift=%d lds=%zu min_blocks=%d: %.4f ms/step\n",
                          c.ntr_shift, c.p_shift, c.lds, c.min_blocks, ms);
            report += line;
            if (ms < best) {
                best = ms;
                chosen = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        timing = tm;
        sched = chosen;
        reset_state();
        sync();
        bool persisted = false;
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            TunedSchedule t;
            t.ntr_shift = chosen.ntr_shift;
            t.p_shift = chosen.p_shift;
            t.lds = chosen.lds;
            t.min_blocks = chosen.min_blocks;
            g_tune[key] = t;
            persisted = tune_write(path, key, t);
        }
        char line[256];
        std::snprintf(line, sizeof(line), "chosen: ntr_shift=%d p_shift=%d lds=%zu min_blocks=%d (%s)\n", chosen.ntr_shift,
                      chosen.p_shift, chosen.lds, chosen.min_blocks,
                      persisted ? ("cached in " + path).c_str() : "cached for this process");
        return report + line;
    }

    // ------------------------------------------------------------------ adjoint (VJP)
    size_t ckpt_slot_bytes() const { return (size_t)(1 + 2 * dim) * g.N * sizeof(T); }
    void set_checkpointing(int every) override {
        KWH_REQUIRE(every >= 0, KWH_EINVAL, "checkpoint interval must be >= 0");
        KWH_REQUIRE(!axisym || every == 0, KWH_EUNSUPPORTED, "the axisymmetric solver has no adjoint");
        sync();
        ckpt_every = every;
        refresh_plane();   // checkpoints and the adjoint run with the split passes
        ckpt_have.clear();
        ckpt.alloc(0);
        if (every > 0) {
            const int n = (nt + every - 1) / every;
            ckpt.alloc((size_t)n * ckpt_slot_bytes());
            ckpt_have.assign(n, 0);
        }
    }
    // state fields in checkpoint order
    std::vector<T*> ckpt_fields() {
        std::vector<T*> f = {p.as<T>()};
        for (int a = 0; a < dim; ++a) f.push_back(u[a].as<T>());
        for (int a = 0; a < dim; ++a) f.push_back(rho[a].as<T>());
        return f;
    }
    void copy_state(char* base, bool to_ckpt) {
        const size_t fb = g.N * sizeof(T);
        int i = 0;
        for (T* f : ckpt_fields()) {
            char* c = base + (size_t)i++ * fb;
            HIPCHK(hipMemcpyAsync(to_ckpt ? (void*)c : (void*)f, to_ckpt ? (const void*)f : (const void*)c, fb,
                                  hipMemcpyDeviceToDevice, stream));
        }
    }
    void save_checkpoint(int k) {
        copy_state(reinterpret_cast<char*>(ckpt.p) + (size_t)k * ckpt_slot_bytes(), true);
        ckpt_have[k] = 1;
    }
    // spec[0] = R2C_z(p), the derived state a step starts from
    void rebuild_P() {
        if (plane) {
            forward_to_middle(p.as<T>(), spec[0].as<C>(), T(1));
            P_valid = true;
            return;
        }
        if (rx[2]) {
            const RegZ z = regz();
            timed(K_ZFR, real_bytes() + spec_bytes(), [&] {
                rx[2]->z_forward_real(em(), dim3(z.blocks), dim3(z.threads), z.lds, (const T*)p.as<T>(),
                                      spec[0].as<C>(), T(1), g, z.P, tab);
            });
        } else {
            ZCfg zc = zcfg(2);
            hipLaunchKernelGGL((z_forward_real<T>), dim3(zc.blocks), dim3(zc.threads), zc.lds, stream, p.as<T>(),
                               spec[0].as<C>(), T(1), g, zc.zt, tab);
            HIPCHK(hipGetLastError());
        }
        P_valid = true;
    }

    // dst (= or +=) sum_a D-_a[scale r_a s_a^spow lu_a]
    // Adjoint row kernels (register-resident path): a diagonal update of the reverse sweep fused
    // into the contiguous-axis transform, followed / preceded by the plain axis-1 passes.
    void adj_forward(const AdjRowArgs<T>& aa, int mode) {
        const RegZ z = regz();
        timed(K_ZFR, real_bytes() * (mode == 0 ? 2 : 4) + spec_bytes(), [&] {
            rx[2]->adj_fwd(em(), dim3(z.blocks), dim3(z.threads), z.lds, aa, g, z.P, tab, mode);
        });
        if (plain1) strided(1, false, {(void*)aa.spec});
    }
    void adj_inverse(C* sp, T* dst, const T* s, int ia, int mode) {
        if (plain1) strided(1, true, {(void*)sp});
        const RegZ z = regz();
        timed(K_ZIR, real_bytes() * (mode == 0 ? 1 : 2) + spec_bytes(), [&] {
            rx[2]->adj_inv(em(), dim3(z.blocks), dim3(z.threads), z.lds, (const C*)sp, dst, s, ia, mode, g, z.P, tab);
        });
    }

    void adj_sum_dneg(T* dst, bool accumulate, const std::vector<DBuf>& lu, int spow, T scale, DBuf& w) {
        std::vector<const void*> ins;
        std::vector<MidOp> ops;
        for (int a = 0; a < dim; ++a) {
            const int ia = a + g.off;
            if (rx[2]) {   // w = scale r_a s_a^spow lu_a fused into the forward transform
                AdjRowArgs<T> aa{};
                aa.lu = lu[a].as<T>();
                aa.r = dtrs[a];
                aa.s = tab.pmlsg[ia];
                aa.ia = ia;
                aa.spow = spow;
                aa.scale = scale;
                aa.spec = spec[a].as<C>();
                adj_forward(aa, 0);
            } else {
                hipLaunchKernelGGL((adj_prediv_kernel<T>), dim3(grid_blocks(g.N)), dim3(256), 0, stream, w.as<T>(),
                                   lu[a].as<T>(), dtrs[a], tab.pmlsg[ia], ia, spow, scale, g);
                HIPCHK(hipGetLastError());
                forward_to_middle(w.as<T>(), spec[a].as<C>(), T(1));
            }
            ins.push_back(spec[a].p);
            ops.push_back(MidOp{OP_DDXN_KAPPA, ia, a, spec[a].p, kappa.p});
        }
        middle(K_MD, ins, ops);
        for (int a = 0; a < dim; ++a) inverse_from_middle(spec[a].as<C>(), dst, T(1), accumulate || a > 0);
    }

    // Gradient of <dL/dy, y> (y = the recorded pressure series) with respect to p0 and c.
    void vjp(const void* dldy, int every, void* grad_p0, void* grad_c0) override {
        KWH_REQUIRE(dldy && grad_p0 && grad_c0, KWH_EINVAL, "NULL argument");
        KWH_REQUIRE(every > 0, KWH_EINVAL, "checkpoint interval must be > 0");
        KWH_REQUIRE(!axisym, KWH_EUNSUPPORTED, "the axisymmetric solver has no adjoint");
        KWH_REQUIRE(has_p0, KWH_EUNSUPPORTED, "the gradient is defined for the p0 problem");
        KWH_REQUIRE(!absorbing && !nonlinear, KWH_EUNSUPPORTED, "the adjoint covers the lossless linear loop");
        KWH_REQUIRE(!psrc.present && !usrc[0].present && !usrc[1].present && !usrc[2].present, KWH_EUNSUPPORTED,
                    "the adjoint does not cover time-varying sources");
        KWH_REQUIRE(!smoothed_media && !smoothed_p0, KWH_EUNSUPPORTED,
                    "the adjoint needs smoothP0 / smoothC0 / smoothRho0 off");
        KWH_REQUIRE(store_p && (record & KWH_REC_P) && ndir == 0, KWH_EINVAL,
                    "the adjoint needs the pressure series recorded (no directivity)");
        HIPCHK(hipSetDevice(dev));
        const int K = every;
        const int nseg = (nt + K - 1) / K;
        bool have = ckpt_every == K && steps == nt && (int)ckpt_have.size() == nseg;
        for (int k = 0; have && k < nseg; ++k) have = ckpt_have[k] != 0;
        if (!have) {   // forward pass with checkpoints
            const bool gr = use_graph;
            set_checkpointing(K);
            reset_state();
            launch(0, nt);
            use_graph = gr;
        }
        sync();
        const int64_t N = g.N;
        // keep the final state (the outputs pFinal / uFinal read it) across the reverse sweep
        DBuf fin;
        fin.alloc(ckpt_slot_bytes());
        copy_state(reinterpret_cast<char*>(fin.p), true);
        // cotangent, time-major on the device
        std::vector<T> h((size_t)nt * nsens);
        const T* src = reinterpret_cast<const T*>(dldy);
        for (int64_t k = 0; k < nsens; ++k)
            for (int t = 0; t < nt; ++t) h[(size_t)t * nsens + k] = src[(size_t)k * nt + t];
        DBuf dg;
        upload(dg, h);
        std::vector<DBuf> lu(dim), lr(dim);
        for (int a = 0; a < dim; ++a) {
            lu[a].alloc(N * sizeof(T));
            lr[a].alloc(N * sizeof(T));
            fill<T>(lu[a].as<T>(), T(0), N);
            fill<T>(lr[a].as<T>(), T(0), N);
        }
        DBuf lp, w, gc, hist, t0;
        lp.alloc(N * sizeof(T));
        w.alloc(N * sizeof(T));
        gc.alloc(N * sizeof(T));
        fill<T>(gc.as<T>(), T(0), N);
        hist.alloc((size_t)K * N * sizeof(T));
        auto scatter = [&](int n) {
            hipLaunchKernelGGL((adj_scatter_kernel<T>), dim3(grid_blocks(nsens * ncorner)), dim3(256), 0, stream,
                               lp.as<T>(), dg.as<T>() + (size_t)n * nsens, sidx.as<int64_t>(),
                               cart ? sw.as<T>() : nullptr, ncorner, nsens);
            HIPCHK(hipGetLastError());
        };
        const size_t fb = N * sizeof(T);
        for (int seg = nseg - 1; seg >= 0; --seg) {
            const int s0 = seg * K, s1 = std::min(s0 + K, nt);
            // recompute the segment's pressures from its checkpoint
            copy_state(reinterpret_cast<char*>(ckpt.p) + (size_t)seg * ckpt_slot_bytes(), false);
            rebuild_P();
            HIPCHK(hipMemcpyAsync(hist.p, p.p, fb, hipMemcpyDeviceToDevice, stream));
            for (int t = s0 + 1; t < s1; ++t) {
                main_step(t);
                HIPCHK(hipMemcpyAsync(reinterpret_cast<char*>(hist.p) + (size_t)(t - s0) * fb, p.p, fb,
                                      hipMemcpyDeviceToDevice, stream));
            }
            for (int n = s1 - 1; n >= std::max(s0, 1); --n) {
                // lp = S^T g^n + sum_a D-_a[r_a s_a lu_a]
                if (n < nt - 1) adj_sum_dneg(lp.as<T>(), false, lu, 1, T(1), w);
                else fill<T>(lp.as<T>(), T(0), N);
                scatter(n);
                hipLaunchKernelGGL((adj_grad_c_kernel<T>), dim3(grid_blocks(N)), dim3(256), 0, stream, gc.as<T>(),
                                   (const T*)lp.as<T>(), (const T*)(hist.as<T>() + (size_t)(n - s0) * N), c0s, N);
                HIPCHK(hipGetLastError());
                // lr_a = c^2 lp + q_a^2 lr_a;  lu_a = s_a^2 lu_a + D+_a[m q_a lr_a]
                std::vector<const void*> ins;
                std::vector<MidOp> ops;
                for (int a = 0; a < dim; ++a) {
                    const int ia = a + g.off;
                    if (rx[2]) {   // lr update and w = m q lr fused into the forward transform
                        AdjRowArgs<T> aa{};
                        aa.lr = lr[a].as<T>();
                        aa.lp = lp.as<T>();
                        aa.c0 = c0s;
                        aa.rho0 = rho0s;
                        aa.s = tab.pml[ia];
                        aa.ia = ia;
                        aa.dt = T(dt);
                        aa.spec = spec[a].as<C>();
                        adj_forward(aa, 1);
                    } else {
                        hipLaunchKernelGGL((adj_rho_kernel<T>), dim3(grid_blocks(N)), dim3(256), 0, stream,
                                           lr[a].as<T>(), w.as<T>(), (const T*)lp.as<T>(), c0s, rho0s, T(dt),
                                           tab.pml[ia], ia, g);
                        HIPCHK(hipGetLastError());
                        forward_to_middle(w.as<T>(), spec[a].as<C>(), T(1));
                    }
                    ins.push_back(spec[a].p);
                    ops.push_back(MidOp{OP_DDXP_KAPPA, ia, a, spec[a].p, kappa.p});
                }
                middle(K_MG, ins, ops);
                for (int a = 0; a < dim; ++a) {
                    const int ia = a + g.off;
                    if (rx[2]) {   // lu = s^2 lu + C2R(...) in one pass
                        adj_inverse(spec[a].as<C>(), lu[a].as<T>(), tab.pmlsg[ia], ia, 2);
                        continue;
                    }
                    hipLaunchKernelGGL((adj_u_decay_kernel<T>), dim3(grid_blocks(N)), dim3(256), 0, stream,
                                       lu[a].as<T>(), tab.pmlsg[ia], ia, g);
                    HIPCHK(hipGetLastError());
                    inverse_from_middle(spec[a].as<C>(), lu[a].as<T>(), T(1), true);
                }
            }
        }
        // n = 0: the p0 override
        if (nt > 1) adj_sum_dneg(lp.as<T>(), false, lu, 1, T(1), w);
        else fill<T>(lp.as<T>(), T(0), N);
        scatter(0);
        t0.alloc(N * sizeof(T));
        if (nt > 1) adj_sum_dneg(t0.as<T>(), false, lu, 2, T(0.5), w);
        else fill<T>(t0.as<T>(), T(0), N);
        DBuf gp0;
        gp0.alloc(N * sizeof(T));
        AdjFinalArgs<T> fa{};
        fa.gp0 = gp0.as<T>();
        fa.gc = gc.as<T>();
        fa.lp0 = lp.as<T>();
        fa.t0 = t0.as<T>();
        for (int a = 0; a < dim; ++a) {
            fa.lr[a] = lr[a].as<T>();
            fa.q[a] = tab.pml[a + g.off];
        }
        fa.p0 = p0s.as<T>();
        fa.c0 = c0s;
        fa.dim = dim;
        hipLaunchKernelGGL((adj_final_kernel<T>), dim3(grid_blocks(N)), dim3(256), 0, stream, fa, g);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grad_p0, gp0.p, fb, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(grad_c0, gc.p, fb, hipMemcpyDeviceToHost, stream));
        // back to the end-of-run state
        copy_state(reinterpret_cast<char*>(fin.p), false);
        rebuild_P();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // combineSensorData on the device (KWaveTransducer.swift:263-289).
    void element_traces(int field, const int32_t* element, const int32_t* delay, int nel, double scale, void* dst,
                        size_t bytes) override {
        const DBuf* ser = nullptr;
        if (field == KWH_OUT_P) {
            KWH_REQUIRE(store_p && (record & KWH_REC_P), KWH_EINVAL, "p not recorded");
            ser = &ser_p;
        } else if (field >= KWH_OUT_UX && field <= KWH_OUT_UZ) {
            const int c = field - KWH_OUT_UX;
            KWH_REQUIRE(c < dim && ser_u[c].p && (record & (KWH_REC_UX << c)), KWH_EINVAL,
                        "velocity series not recorded");
            ser = &ser_u[c];
        } else {
            KWH_REQUIRE(false, KWH_EINVAL, "element traces need a recorded p or u series");
        }
        KWH_REQUIRE(element && delay && nel > 0, KWH_EINVAL, "element / delay arrays missing");
        KWH_REQUIRE(bytes >= (size_t)nel * rec_steps * sizeof(T), KWH_EINVAL, "destination buffer too small");
        std::vector<int> rowptr(nel + 1, 0), rows(nsens), dl(nsens);
        for (int64_t k = 0; k < nsens; ++k) {
            KWH_REQUIRE(element[k] >= 0 && element[k] < nel, KWH_EINVAL, "element index out of range");
            KWH_REQUIRE(delay[k] >= 0, KWH_EINVAL, "receive delays must be >= 0");
            rowptr[element[k] + 1]++;
            dl[k] = delay[k];
        }
        for (int e = 0; e < nel; ++e) rowptr[e + 1] += rowptr[e];
        std::vector<int> fillp(rowptr.begin(), rowptr.end() - 1);
        for (int64_t k = 0; k < nsens; ++k) rows[fillp[element[k]]++] = (int)k;
        HIPCHK(hipSetDevice(dev));
        sync();
        DBuf drp, drows, ddl, out;
        upload(drp, rowptr);
        upload(drows, rows);
        upload(ddl, dl);
        const int64_t total = (int64_t)nel * rec_steps;
        out.alloc(std::max<int64_t>(total, 1) * sizeof(T));
        if (total > 0) {
            hipLaunchKernelGGL((element_traces_kernel<T>), dim3(grid_blocks(total)), dim3(256), 0, stream, ser->as<T>(),
                               nsens, rec_steps, drp.as<int>(), drows.as<int>(), ddl.as<int>(), nel, scale, out.as<T>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(dst, out.p, total * sizeof(T), hipMemcpyDeviceToHost, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
    }

    // ------------------------------------------------------------------ API
    void check_range(int t0, int t1) {
        KWH_REQUIRE(t0 == steps, KWH_ESTATE, "runs must be contiguous: t_begin must equal the steps already run");
        KWH_REQUIRE(t1 >= t0 && t1 <= nt, KWH_EINVAL, "t_end must be in [t_begin, nt]");
    }

    void launch(int t0, int t1) override {
        check_range(t0, t1);
        HIPCHK(hipSetDevice(dev));
        hipLaunchKernelGGL(set_counters_kernel, dim3(1), dim3(64), 0, stream, (int*)ctr.p, t0, rec_steps);
        HIPCHK(hipGetLastError());
        for (int t = t0; t < t1; ++t) step(t);
    }
    void sync() override {
        HIPCHK(hipStreamSynchronize(stream));
        if (timing) collect_timing();
    }
    void set_nonfinite_check(int every) override {
        KWH_REQUIRE(every >= 0, KWH_EINVAL, "check interval must be >= 0");
        nonfinite_every = every;
        if (every > 0 && !nf_flag.p) nf_flag.alloc(sizeof(int));
    }
    // Scan p after step t (enqueued; the flag is read by nonfinite_seen after a sync).
    void nonfinite_scan(int t) {
        hipLaunchKernelGGL((nonfinite_kernel<T>), dim3(grid_blocks(g.N)), dim3(256), 0, stream, p.as<T>(), g.N, t,
                           nf_flag.as<int>());
        HIPCHK(hipGetLastError());
    }
    // After a sync: throw KWH_ENONFINITE if a scan flagged a step.
    void nonfinite_raise() {
        int first = INT_MAX;
        HIPCHK(hipMemcpy(&first, nf_flag.p, sizeof(int), hipMemcpyDeviceToHost));
        if (first != INT_MAX) {
            char msg[160];
            std::snprintf(msg, sizeof(msg),
                          "non-finite pressure after step %d (checked every %d steps); the run stopped after step %d",
                          first, nonfinite_every, steps - 1);
            throw KwhError(KWH_ENONFINITE, msg);
        }
    }
    void run(int t0, int t1, kwh_step_cb cb, int interval, void* user) override {
        check_range(t0, t1);
        const int ck = nonfinite_every;
        if (ck > 0) {
            HIPCHK(hipSetDevice(dev));
            hipLaunchKernelGGL((fill_kernel<int>), dim3(1), dim3(64), 0, stream, nf_flag.as<int>(), INT_MAX, (int64_t)1);
            HIPCHK(hipGetLastError());
        }
        // step t is checked when (t + 1) % ck == 0 and after the last step of the call
        auto check_due = [&](int t) { return ck > 0 && (((t + 1) % ck) == 0 || t + 1 == t1); };
        if (!cb) {
            if (ck <= 0) {
                launch(t0, t1);
                sync();
                return;
            }
            // chunks of ck steps, each followed by a scan and a read of the flag: the run stops at
            // the first chunk whose end state is not finite
            for (int t = t0; t < t1;) {
                const int te = std::min(t1, (t / ck + 1) * ck);
                launch(t, te);
                nonfinite_scan(te - 1);
                sync();
                nonfinite_raise();
                t = te;
            }
            return;
        }
        // Steps run in chunks of up to kCallbackChunk with one sync each; every step still gets its
        // callback, in order, after its chunk. A chunk ends at the next step whose pressure the
        // callback receives (t % interval == 0; interval <= 0: none) and at the check interval.
        constexpr int kCallbackChunk = 64;
        const bool want_p = interval > 0;
        const int iv = std::max(interval, 1);
        std::vector<T> host;
        for (int t = t0; t < t1;) {
            int te = std::min(t1, t + kCallbackChunk);
            if (want_p) {
                const int next_due = ((t + iv - 1) / iv) * iv;
                if (next_due < te) te = next_due + 1;
            }
            if (ck > 0) te = std::min(te, (t / ck + 1) * ck);
            launch(t, te);
            const int last = te - 1;
            const bool due = want_p && (last % iv) == 0;
            if (due) {
                host.resize(g.N);
                HIPCHK(hipMemcpyAsync(host.data(), p.p, g.N * sizeof(T), hipMemcpyDeviceToHost, stream));
            }
            const bool scan = check_due(last);
            if (scan) nonfinite_scan(last);
            sync();
            for (int s = t; s <= last; ++s) cb(s, nt, (s == last && due) ? host.data() : nullptr, user);
            if (scan) nonfinite_raise();
            t = te;
        }
    }

    int steps_done() const override { return steps; }
    int elem_bytes() const override { return (int)sizeof(T); }

    void output_shape(int field, int64_t shape[3], int* ndim) override {
        shape[0] = shape[1] = shape[2] = 1;
        auto grid = [&] { *ndim = dim; for (int a = 0; a < dim; ++a) shape[a] = nphys[a]; };
        auto series = [&] { *ndim = 2; shape[0] = nsens; shape[1] = rec_steps; };
        auto agg = [&] { *ndim = 1; shape[0] = nsens; };
        const int uc = (field >= KWH_OUT_UX && field <= KWH_OUT_IZ_AVG) ? ((field - KWH_OUT_UX) % 3) : 0;
        if (nsens == 0 && field != KWH_OUT_P_FINAL && !(field >= KWH_OUT_UX_FINAL && field <= KWH_OUT_UZ_FINAL)) {
            // a sensor mask without points: every recorded sensor field is empty ([0, nt] series,
            // [0] aggregates), as the reference's gather over no indices
            uint32_t bit = 0;
            bool ser = false;
            switch (field) {
                case KWH_OUT_P: bit = KWH_REC_P; ser = true; break;
                case KWH_OUT_P_MAX: bit = KWH_REC_P_MAX; break;
                case KWH_OUT_P_MIN: bit = KWH_REC_P_MIN; break;
                case KWH_OUT_P_RMS: bit = KWH_REC_P_RMS; break;
                case KWH_OUT_UX: case KWH_OUT_UY: case KWH_OUT_UZ: bit = (uint32_t)KWH_REC_UX << uc; ser = true; break;
                case KWH_OUT_UX_MAX: case KWH_OUT_UY_MAX: case KWH_OUT_UZ_MAX: bit = KWH_REC_U_MAX; break;
                case KWH_OUT_UX_RMS: case KWH_OUT_UY_RMS: case KWH_OUT_UZ_RMS: bit = KWH_REC_U_RMS; break;
                case KWH_OUT_IX_AVG: case KWH_OUT_IY_AVG: case KWH_OUT_IZ_AVG: bit = KWH_REC_I_AVG; break;
                default: throw KwhError(KWH_EINVAL, "unknown output field");
            }
            KWH_REQUIRE((record & bit) && uc < dim, KWH_EINVAL, "field not recorded");
            *ndim = ser ? 2 : 1;
            shape[0] = 0;
            if (ser) shape[1] = steps;   // no step is recorded without points; the series has nt columns
            return;
        }
        switch (field) {
            case KWH_OUT_P_FINAL: grid(); return;
            case KWH_OUT_P: KWH_REQUIRE(store_p && (record & KWH_REC_P), KWH_EINVAL, "p not recorded"); series(); return;
            case KWH_OUT_P_MAX: KWH_REQUIRE(pmax.p, KWH_EINVAL, "pMax not recorded"); agg(); return;
            case KWH_OUT_P_MIN: KWH_REQUIRE(pmin.p, KWH_EINVAL, "pMin not recorded"); agg(); return;
            case KWH_OUT_P_RMS: KWH_REQUIRE(psum2.p, KWH_EINVAL, "pRms not recorded"); agg(); return;
            case KWH_OUT_UX: case KWH_OUT_UY: case KWH_OUT_UZ:
                KWH_REQUIRE(uc < dim && ser_u[uc].p && (record & (KWH_REC_UX << uc)), KWH_EINVAL, "velocity series not recorded");
                series(); return;
            case KWH_OUT_UX_MAX: case KWH_OUT_UY_MAX: case KWH_OUT_UZ_MAX: {
                const int c = field - KWH_OUT_UX_MAX;
                KWH_REQUIRE(c < dim && umax[c].p, KWH_EINVAL, "uMax not recorded"); agg(); return; }
            case KWH_OUT_UX_RMS: case KWH_OUT_UY_RMS: case KWH_OUT_UZ_RMS: {
                const int c = field - KWH_OUT_UX_RMS;
                KWH_REQUIRE(c < dim && usum2[c].p, KWH_EINVAL, "uRms not recorded"); agg(); return; }
            case KWH_OUT_UX_FINAL: case KWH_OUT_UY_FINAL: case KWH_OUT_UZ_FINAL: {
                const int c = field - KWH_OUT_UX_FINAL;
                KWH_REQUIRE(c < dim && (record & KWH_REC_U_FINAL), KWH_EINVAL, "uFinal not recorded"); grid(); return; }
            case KWH_OUT_IX_AVG: case KWH_OUT_IY_AVG: case KWH_OUT_IZ_AVG: {
                const int c = field - KWH_OUT_IX_AVG;
                KWH_REQUIRE(c < dim && (record & KWH_REC_I_AVG) && ser_p.p && ser_u[c].p, KWH_EINVAL, "iAvg not recorded");
                agg(); return; }
            default: throw KwhError(KWH_EINVAL, "unknown output field");
        }
    }

    void read_output(int field, void* dst, size_t bytes) override {
        int64_t shape[3];
        int nd = 0;
        output_shape(field, shape, &nd);
        int64_t count = 1;
        for (int i = 0; i < nd; ++i) count *= shape[i];
        KWH_REQUIRE(bytes >= (size_t)count * sizeof(T), KWH_EINVAL, "destination buffer too small");
        if (count == 0) return;   // empty sensor fields
        HIPCHK(hipSetDevice(dev));
        sync();
        auto d2h = [&](const void* src, size_t b) {
            HIPCHK(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        };
        auto series_out = [&](const DBuf& s) {
            std::vector<T> h((size_t)rec_steps * nsens);
            HIPCHK(hipMemcpyAsync(h.data(), s.p, h.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            T* o = reinterpret_cast<T*>(dst);
            for (int64_t s_ = 0; s_ < nsens; ++s_)
                for (int t = 0; t < rec_steps; ++t) o[s_ * rec_steps + t] = h[(size_t)t * nsens + s_];
        };
        auto rms_out = [&](const DBuf& s2) {
            DBuf tmp;
            tmp.alloc(nsens * sizeof(T));
            hipLaunchKernelGGL((rms_kernel<T>), dim3(grid_blocks(nsens)), dim3(256), 0, stream, s2.as<double>(), tmp.as<T>(),
                               nsens, 1.0 / std::max(rec_steps, 1));
            HIPCHK(hipGetLastError());
            d2h(tmp.p, nsens * sizeof(T));
        };
        switch (field) {
            case KWH_OUT_P_FINAL: d2h(p.p, g.N * sizeof(T)); return;
            case KWH_OUT_P: series_out(ser_p); return;
            case KWH_OUT_P_MAX: d2h(pmax.p, nsens * sizeof(T)); return;
            case KWH_OUT_P_MIN: d2h(pmin.p, nsens * sizeof(T)); return;
            case KWH_OUT_P_RMS: rms_out(psum2); return;
            case KWH_OUT_UX: case KWH_OUT_UY: case KWH_OUT_UZ: series_out(ser_u[field - KWH_OUT_UX]); return;
            case KWH_OUT_UX_MAX: case KWH_OUT_UY_MAX: case KWH_OUT_UZ_MAX: d2h(umax[field - KWH_OUT_UX_MAX].p, nsens * sizeof(T)); return;
            case KWH_OUT_UX_RMS: case KWH_OUT_UY_RMS: case KWH_OUT_UZ_RMS: rms_out(usum2[field - KWH_OUT_UX_RMS]); return;
            case KWH_OUT_UX_FINAL: case KWH_OUT_UY_FINAL: case KWH_OUT_UZ_FINAL: d2h(u[field - KWH_OUT_UX_FINAL].p, g.N * sizeof(T)); return;
            case KWH_OUT_IX_AVG: case KWH_OUT_IY_AVG: case KWH_OUT_IZ_AVG: {
                const int c = field - KWH_OUT_IX_AVG;
                DBuf tmp;
                tmp.alloc(nsens * sizeof(T));
                hipLaunchKernelGGL((intensity_kernel<T>), dim3(grid_blocks(nsens, 64)), dim3(64), 0, stream, ser_p.as<T>(),
                                   ser_u[c].as<T>(), tmp.as<T>(), nsens, rec_steps);
                HIPCHK(hipGetLastError());
                d2h(tmp.p, nsens * sizeof(T));
                return;
            }
            default: throw KwhError(KWH_EINVAL, "unknown output field");
        }
    }

    // read_output without device allocations: every staging buffer lives in the caller-provided
    // device scratch, and the series and per-sensor outputs are produced in chunks of sensors that
    // fit it. Same kernels and copies as read_output, so the results are identical.
    void read_output_noalloc(int field, void* dst, size_t bytes, void* scr, size_t scr_bytes) override {
        int64_t shape[3];
        int nd = 0;
        output_shape(field, shape, &nd);
        int64_t count = 1;
        for (int i = 0; i < nd; ++i) count *= shape[i];
        KWH_REQUIRE(bytes >= (size_t)count * sizeof(T), KWH_EINVAL, "destination buffer too small");
        if (count == 0) return;
        const bool series = field == KWH_OUT_P || (field >= KWH_OUT_UX && field <= KWH_OUT_UZ);
        KWH_REQUIRE(!series || (size_t)rec_steps * sizeof(T) <= scr_bytes, KWH_EUNSUPPORTED,
                    "recorded series longer than the allocation-free scratch holds (8 MiB per sensor row)");
        HIPCHK(hipSetDevice(dev));
        sync();
        char* out = reinterpret_cast<char*>(dst);
        T* st = reinterpret_cast<T*>(scr);
        auto d2h = [&](void* to, const void* from, size_t b) {
            HIPCHK(hipMemcpyAsync(to, from, b, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        };
        // sensors per chunk when each needs `per` elements of staging
        auto chunks = [&](int64_t per, auto&& body) {
            const int64_t cap = std::max<int64_t>((int64_t)(scr_bytes / (sizeof(T) * (size_t)per)), 1);
            for (int64_t s0 = 0; s0 < nsens; s0 += cap) {
                const int64_t cnt = std::min(cap, nsens - s0);
                body(s0, cnt);
                HIPCHK(hipGetLastError());
                d2h(out + (size_t)s0 * per * sizeof(T), st, (size_t)cnt * per * sizeof(T));
            }
        };
        auto series_rows = [&](const DBuf& ser) {
            chunks(rec_steps, [&](int64_t s0, int64_t cnt) {
                hipLaunchKernelGGL((series_rows_kernel<T>), dim3(grid_blocks(cnt * rec_steps)), dim3(256), 0, stream,
                                   ser.as<T>(), nsens, rec_steps, s0, cnt, st);
            });
        };
        auto rms = [&](const DBuf& s2) {
            chunks(1, [&](int64_t s0, int64_t cnt) {
                hipLaunchKernelGGL((rms_kernel<T>), dim3(grid_blocks(cnt)), dim3(256), 0, stream, s2.as<double>() + s0, st,
                                   cnt, 1.0 / std::max(rec_steps, 1));
            });
        };
        switch (field) {
            case KWH_OUT_P: series_rows(ser_p); return;
            case KWH_OUT_UX: case KWH_OUT_UY: case KWH_OUT_UZ: series_rows(ser_u[field - KWH_OUT_UX]); return;
            case KWH_OUT_P_RMS: rms(psum2); return;
            case KWH_OUT_UX_RMS: case KWH_OUT_UY_RMS: case KWH_OUT_UZ_RMS: rms(usum2[field - KWH_OUT_UX_RMS]); return;
            case KWH_OUT_IX_AVG: case KWH_OUT_IY_AVG: case KWH_OUT_IZ_AVG: {
                const int c = field - KWH_OUT_IX_AVG;
                chunks(1, [&](int64_t s0, int64_t cnt) {
                    hipLaunchKernelGGL((intensity_range_kernel<T>), dim3(grid_blocks(cnt, 64)), dim3(64), 0, stream,
                                       ser_p.as<T>(), ser_u[c].as<T>(), st, nsens, rec_steps, s0, cnt);
                });
                return;
            }
            case KWH_OUT_P_FINAL: d2h(dst, p.p, g.N * sizeof(T)); return;
            case KWH_OUT_P_MAX: d2h(dst, pmax.p, nsens * sizeof(T)); return;
            case KWH_OUT_P_MIN: d2h(dst, pmin.p, nsens * sizeof(T)); return;
            case KWH_OUT_UX_MAX: case KWH_OUT_UY_MAX: case KWH_OUT_UZ_MAX:
                d2h(dst, umax[field - KWH_OUT_UX_MAX].p, nsens * sizeof(T));
                return;
            case KWH_OUT_UX_FINAL: case KWH_OUT_UY_FINAL: case KWH_OUT_UZ_FINAL:
                d2h(dst, u[field - KWH_OUT_UX_FINAL].p, g.N * sizeof(T));
                return;
            default: throw KwhError(KWH_EINVAL, "unknown output field");
        }
    }

    // State = p, u, rho, sensor accumulators and series, counters.
    std::vector<std::pair<void*, size_t>> state_parts() {
        std::vector<std::pair<void*, size_t>> v;
        auto add = [&](DBuf& b) { if (b.p) v.push_back({b.p, b.bytes}); };
        add(p);
        for (int a = 0; a < dim; ++a) { add(u[a]); add(rho[a]); }
        add(ser_p); add(pmax); add(pmin); add(psum2);
        for (int a = 0; a < dim; ++a) { add(ser_u[a]); add(umax[a]); add(usum2[a]); }
        return v;
    }
    // A state blob opens with a header that names the sim it belongs to (precision, dimension,
    // grid, steps of the run, sensor points, recorded fields, payload size) and the counters.
    // load_state rejects a blob whose header does not match this sim, or whose counters this sim
    // could not have reached, before anything is copied to the device.
    struct StateHeader {
        uint32_t magic;
        int32_t elem_bytes, dim, nt;
        int32_t n[3];
        uint32_t record;
        int64_t nsens;
        uint64_t payload;
        int32_t steps, rec_steps;
    };
    static constexpr uint32_t kStateMagic = 0x4b574853u;   // "KWHS"
    StateHeader state_header() {
        StateHeader h{};
        h.magic = kStateMagic;
        h.elem_bytes = (int32_t)sizeof(T);
        h.dim = dim;
        h.nt = nt;
        for (int a = 0; a < 3; ++a) h.n[a] = nphys[a];
        h.record = record;
        h.nsens = nsens;
        for (auto& pr : state_parts()) h.payload += pr.second;
        h.steps = steps;
        h.rec_steps = rec_steps;
        return h;
    }
    size_t state_bytes() override { return sizeof(StateHeader) + state_header().payload; }
    void save_state(void* dst, size_t bytes) override {
        KWH_REQUIRE(dst != nullptr, KWH_EINVAL, "state buffer is NULL");
        KWH_REQUIRE(bytes >= state_bytes(), KWH_EINVAL, "state buffer too small");
        sync();
        char* o = reinterpret_cast<char*>(dst);
        const StateHeader hdr = state_header();
        std::memcpy(o, &hdr, sizeof(hdr));
        o += sizeof(hdr);
        for (auto& pr : state_parts()) {
            HIPCHK(hipMemcpyAsync(o, pr.first, pr.second, hipMemcpyDeviceToHost, stream));
            o += pr.second;
        }
        HIPCHK(hipStreamSynchronize(stream));
    }
    void load_state(const void* src, size_t bytes) override {
        KWH_REQUIRE(src != nullptr, KWH_EINVAL, "state buffer is NULL");
        KWH_REQUIRE(bytes >= sizeof(StateHeader), KWH_EINVAL, "state buffer too small");
        StateHeader hdr;
        std::memcpy(&hdr, src, sizeof(hdr));
        const StateHeader mine = state_header();
        KWH_REQUIRE(hdr.magic == kStateMagic, KWH_EINVAL, "not a kwavehip state blob");
        KWH_REQUIRE(hdr.elem_bytes == mine.elem_bytes && hdr.dim == mine.dim && hdr.nt == mine.nt &&
                        hdr.n[0] == mine.n[0] && hdr.n[1] == mine.n[1] && hdr.n[2] == mine.n[2] &&
                        hdr.record == mine.record && hdr.nsens == mine.nsens && hdr.payload == mine.payload,
                    KWH_EINVAL, "state blob belongs to a different simulation (precision, grid, nt, sensor or record)");
        KWH_REQUIRE(bytes >= sizeof(StateHeader) + mine.payload, KWH_EINVAL, "state buffer too small");
        KWH_REQUIRE(hdr.steps >= 0 && hdr.steps <= nt, KWH_EINVAL, "state step counter out of range");
        KWH_REQUIRE(hdr.rec_steps == ((rec_p || rec_u) ? hdr.steps : 0), KWH_EINVAL,
                    "state record counter inconsistent with the step counter");
        sync();
        const char* o = reinterpret_cast<const char*>(src) + sizeof(hdr);
        for (auto& pr : state_parts()) {
            HIPCHK(hipMemcpyAsync(pr.first, o, pr.second, hipMemcpyHostToDevice, stream));
            o += pr.second;
        }
        steps = hdr.steps;
        rec_steps = hdr.rec_steps;
        // P = R2C_z(p) is derived state, rebuilt with the kernel the time step uses (bitwise the
        // spectrum an uninterrupted run carries)
        rebuild_P();
        HIPCHK(hipStreamSynchronize(stream));
    }

    void set_timing(bool on) override {
        sync();
        timing = on;
        timing_only = -1;
        for (int i = 0; i < kNames; ++i) { stat_launch[i] = 0; stat_ms[i] = 0; stat_bytes[i] = 0; }
    }
    void set_timing_only(const char* name) override {
        int id = -1;
        for (int i = 0; i < kNames; ++i)
            if (std::strcmp(kKernelNames[i], name) == 0) id = i;
        KWH_REQUIRE(id >= 0, KWH_EINVAL, "unknown kernel name");
        set_timing(true);
        timing_only = id;
    }
    int kernel_stats(kwh_kernel_stat* st, int cap) override {
        sync();
        int n = 0;
        for (int i = 0; i < kNames; ++i) {
            if (!stat_launch[i]) continue;
            if (n < cap) {
                std::snprintf(st[n].name, sizeof(st[n].name), "%s", kKernelNames[i]);
                st[n].launches = stat_launch[i];
                st[n].total_ms = stat_ms[i];
                st[n].bytes_per_launch = stat_bytes[i] / (double)stat_launch[i];
            }
            ++n;
        }
        return n;
    }
    void set_graph(bool on) override {
        KWH_REQUIRE(!axisym || !on, KWH_EUNSUPPORTED, "the axisymmetric solver runs eagerly (no step graphs)");
        sync();
        use_graph = on;
    }
};

}  // namespace

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------

template <typename F>
static int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return KWH_OK;
    } catch (const KwhError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return KWH_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return KWH_EDEVICE;
    }
}

// Independent simulations advanced together. Three ways:
//  * streams (mode 0): each sim enqueues its whole run eagerly on its own stream; the runtime
//    spreads the streams over the hardware queues, so kernels of small grids overlap;
//  * per-sim graphs (KWH_BATCH_GRAPH): every sim replays its own captured step graphs (one per
//    step key), interleaved step by step;
//  * fused (KWH_BATCH_FUSED): one launch per pass of the step for all sims whose steps have the
//    same launch sequence and launch geometry (kwh_batch.h) -- in practice the sims of one grid
//    shape. Each sim's step is recorded once per step key; the batch concatenates the sims'
//    argument packs per pass, uploads them with a block-owner table, and issues the step as a
//    short list of launches on the batch's stream (the groups of one step one after another),
//    captured in a hipGraph and replayed with KWH_BATCH_GRAPH. Results are bitwise
//    those of the single runs: every workgroup runs the same body on the same tile.
struct kwh_batch {
    int dev = 0;
    std::vector<kwh_sim*> sims;
    std::vector<hipStream_t> streams;   // fused runs: the batch's stream

    struct FusedLaunch {
        const void* fn = nullptr;
        unsigned nblocks = 0, block = 0;
        size_t lds = 0;
        DBuf packs, owner;
    };
    struct FusedStep {
        std::vector<std::unique_ptr<FusedLaunch>> launches;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    // keyed by the members of a group and their step keys
    std::map<std::vector<uint64_t>, std::unique_ptr<FusedStep>> fused;

    ~kwh_batch() {
        for (auto& e : fused) {
            if (e.second->exec) (void)hipGraphExecDestroy(e.second->exec);
            if (e.second->graph) (void)hipGraphDestroy(e.second->graph);
        }
        fused.clear();
        for (hipStream_t st : streams) (void)hipStreamDestroy(st);
        for (kwh_sim* s : sims) delete s;
    }

    // Pack position j of every member's step into one launch.
    std::unique_ptr<FusedStep> build(const std::vector<const StepRecord*>& recs) {
        auto fs = std::make_unique<FusedStep>();
        const size_t nl = recs[0]->launches.size();
        for (const StepRecord* r : recs)
            KWH_REQUIRE(r->launches.size() == nl, KWH_EUNSUPPORTED, "batched step: launch sequences differ");
        for (size_t j = 0; j < nl; ++j) {
            auto fl = std::make_unique<FusedLaunch>();
            const LaunchRec& l0 = recs[0]->launches[j];
            fl->fn = l0.fn;
            const size_t psz = l0.pack.size();
            std::vector<unsigned char> packs;
            std::vector<int> owner;
            for (size_t m = 0; m < recs.size(); ++m) {
                const LaunchRec& l = recs[m]->launches[j];
                KWH_REQUIRE(l.fn == l0.fn && l.pack.size() == psz && (l.flex || l.block == l0.block), KWH_EUNSUPPORTED,
                            "batched step: kernels differ");
                const int start = (int)owner.size();
                const int64_t nb = (int64_t)l.gx * l.gy;
                KWH_REQUIRE((int64_t)owner.size() + nb < (int64_t)1 << 31, KWH_EUNSUPPORTED, "batched launch too large");
                owner.insert(owner.end(), (size_t)nb, (int)m);
                const size_t off = packs.size();
                packs.insert(packs.end(), l.pack.begin(), l.pack.end());
                std::memcpy(packs.data() + off, &start, sizeof(int));   // BatchPack::start
                fl->block = std::max(fl->block, l.block);
                fl->lds = std::max(fl->lds, l.lds);
            }
            fl->nblocks = (unsigned)owner.size();
            if (fl->nblocks == 0) continue;
            if (fl->lds > 65536)   // plane kernels: raise the batched kernel's dynamic-LDS limit
                HIPCHK(hipFuncSetAttribute(fl->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fl->lds));
            fl->packs.alloc(packs.size());
            HIPCHK(hipMemcpy(fl->packs.p, packs.data(), packs.size(), hipMemcpyHostToDevice));
            fl->owner.alloc(owner.size() * sizeof(int));
            HIPCHK(hipMemcpy(fl->owner.p, owner.data(), owner.size() * sizeof(int), hipMemcpyHostToDevice));
            fs->launches.push_back(std::move(fl));
        }
        return fs;
    }

    void launch_fused(FusedStep& fs, hipStream_t st) {
        for (auto& fl : fs.launches) {
            void* pk = fl->packs.p;
            void* ow = fl->owner.p;
            void* args[] = {&pk, &ow};
            HIPCHK(hipLaunchKernel(fl->fn, dim3(fl->nblocks), dim3(fl->block), args, fl->lds, st));
        }
    }

    void run_fused(bool use_graph) {
        const int n = (int)sims.size();
        // Every grid shape has a stream of its own (a sim's steps always go to its shape's stream,
        // so they stay ordered) and the shapes' batched launches overlap, filling each other's
        // launch gaps: 48 scan lines in plane mode 23.5-24.3 -> 27.2 Gcell-steps/s, with the split
        // passes 15.9-16.2 -> 17.7. KWH_BATCH_STREAMS=0: one stream, the shapes one after another.
        const char* es = std::getenv("KWH_BATCH_STREAMS");
        const bool per_shape = !es || std::atoi(es) != 0;
        std::map<uint64_t, int> shape_stream;
        std::vector<int> sim_stream(n, 0);
        if (per_shape)
            for (int i = 0; i < n; ++i) {
                auto it = shape_stream.find(sims[i]->shape_class());
                if (it == shape_stream.end())
                    it = shape_stream.emplace(sims[i]->shape_class(), (int)shape_stream.size()).first;
                sim_stream[i] = it->second;
            }
        const int nst = per_shape ? std::max<int>(1, (int)shape_stream.size()) : 1;
        while ((int)streams.size() < nst) {
            hipStream_t s = nullptr;
            HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            streams.push_back(s);
        }
        for (int i = 0; i < n; ++i) sims[i]->set_counters_on(streams[sim_stream[i]]);
        std::vector<const StepRecord*> rec(n);
        std::vector<int> ts(n);
        for (;;) {
            // group the active sims by the launch sequence of their current step
            std::vector<std::pair<uint64_t, int>> order;
            for (int i = 0; i < n; ++i) {
                ts[i] = sims[i]->steps_done();
                if (ts[i] >= sims[i]->total_steps()) continue;
                rec[i] = &sims[i]->recorded_step(ts[i]);
                order.push_back({rec[i]->signature, i});
            }
            if (order.empty()) break;
            std::stable_sort(order.begin(), order.end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            for (size_t g0 = 0; g0 < order.size();) {
                size_t g1 = g0;
                std::vector<uint64_t> key;
                std::vector<const StepRecord*> members;
                while (g1 < order.size() && order[g1].first == order[g0].first) {
                    const int i = order[g1].second;
                    key.push_back((uint64_t)i);
                    key.push_back(rec[i]->key);
                    members.push_back(rec[i]);
                    ++g1;
                }
                auto it = fused.find(key);
                if (it == fused.end()) it = fused.emplace(key, build(members)).first;
                FusedStep& fs = *it->second;
                hipStream_t st = streams[sim_stream[order[g0].second]];
                if (use_graph) {
                    if (!fs.exec) {
                        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                        launch_fused(fs, st);
                        HIPCHK(hipStreamEndCapture(st, &fs.graph));
                        HIPCHK(hipGraphInstantiate(&fs.exec, fs.graph, nullptr, nullptr, 0));
                    }
                    HIPCHK(hipGraphLaunch(fs.exec, st));
                } else {
                    launch_fused(fs, st);
                }
                g0 = g1;
            }
            for (auto& o : order) sims[o.second]->step_done(ts[o.second]);
        }
        for (int k = 0; k < nst; ++k) HIPCHK(hipStreamSynchronize(streams[k]));
    }

    void run(int mode) {
        HIPCHK(hipSetDevice(dev));
        for (kwh_sim* s : sims) s->sync();
        if (mode & KWH_BATCH_PLANE)
            for (kwh_sim* s : sims)
                if (s->plane_available()) s->set_plane(true);
        if (mode & KWH_BATCH_FUSED) {
            run_fused((mode & KWH_BATCH_GRAPH) != 0);
            return;
        }
        if (!(mode & KWH_BATCH_GRAPH)) {
            for (kwh_sim* s : sims) s->launch(s->steps_done(), s->total_steps());
            for (kwh_sim* s : sims) s->sync();
            return;
        }
        for (kwh_sim* s : sims) s->set_counters_on(nullptr);
        for (bool any = true; any;) {
            any = false;
            for (kwh_sim* s : sims) {
                const int t = s->steps_done();
                if (t >= s->total_steps()) continue;
                s->graph_step(t);
                any = true;
            }
        }
        for (kwh_sim* s : sims) s->sync();
    }
};

static void validate_desc(const kwh_sim_desc* d) {
    KWH_REQUIRE(d != nullptr, KWH_EINVAL, "descriptor is NULL");
    KWH_REQUIRE(d->abi_version == KWH_ABI_VERSION, KWH_EINVAL, "descriptor ABI version mismatch");
    KWH_REQUIRE(d->dim >= 1 && d->dim <= 3, KWH_EUNSUPPORTED, "kspaceFirstOrder: dim must be 1, 2 or 3");
    for (int a = 0; a < d->dim; ++a) {
        KWH_REQUIRE(d->n[a] >= 1, KWH_EINVAL, "grid sizes must be >= 1");
        KWH_REQUIRE(d->d[a] > 0 && std::isfinite(d->d[a]), KWH_EINVAL, "grid spacing must be positive");
    }
    KWH_REQUIRE(d->nt > 0, KWH_EINVAL, "call grid.makeTime(...) before running the solver");
    KWH_REQUIRE(d->dt > 0 && std::isfinite(d->dt), KWH_EINVAL, "dt must be positive");
    KWH_REQUIRE(d->dtype == KWH_FLOAT32 || d->dtype == KWH_FLOAT64, KWH_EINVAL, "dtype must be float32 or float64");
    KWH_REQUIRE(d->sound_speed.data && d->sound_speed.count >= 1, KWH_EINVAL, "soundSpeed is required");
    KWH_REQUIRE(d->density.data && d->density.count >= 1, KWH_EINVAL, "density is required");
    for (int64_t i = 0; i < d->sound_speed.count; ++i)
        KWH_REQUIRE(d->sound_speed.data[i] > 0 && std::isfinite(d->sound_speed.data[i]), KWH_EINVAL,
                    "soundSpeed must be positive and finite");
    for (int64_t i = 0; i < d->density.count; ++i)
        KWH_REQUIRE(d->density.data[i] > 0 && std::isfinite(d->density.data[i]), KWH_EINVAL,
                    "density must be positive and finite");
    KWH_REQUIRE(d->alpha_mode >= 0 && d->alpha_mode <= 3, KWH_EINVAL, "invalid alpha mode");
    KWH_REQUIRE(d->pml_size >= 0, KWH_EINVAL, "pmlSize must be >= 0");
    KWH_REQUIRE(std::isfinite(d->pml_alpha), KWH_EINVAL, "pmlAlpha must be finite");
    int64_t N = 1;
    for (int a = 0; a < d->dim; ++a) {
        N *= d->n[a];
        KWH_REQUIRE(d->pml_size <= d->n[a], KWH_EINVAL, "pmlSize must not exceed the grid size");
    }
    auto field_ok = [&](const kwh_field& f, const char* name) {
        if (!f.data) return;
        KWH_REQUIRE(f.count == 1 || f.count == N, KWH_EINVAL, std::string(name) + " must be scalar or grid-shaped");
    };
    field_ok(d->sound_speed, "soundSpeed");
    field_ok(d->density, "density");
    field_ok(d->alpha_coeff, "alphaCoeff");
    field_ok(d->bona, "bOnA");
    if (d->alpha_coeff.data && d->alpha_mode != KWH_ALPHA_NONE && d->alpha_mode != KWH_ALPHA_STOKES)
        KWH_REQUIRE(std::isfinite(d->alpha_power), KWH_EINVAL, "alphaPower must be finite");
    if (d->p0.data) {
        KWH_REQUIRE(d->p0.count == N, KWH_EINVAL, "p0 must be grid-shaped");
        for (int64_t i = 0; i < N; ++i)
            KWH_REQUIRE(std::isfinite(d->p0.data[i]), KWH_EINVAL, "p0 contains non-finite values");
    }
    auto src_ok = [&](const kwh_source_field& s, const char* name) {
        if (!s.index || !s.signal || s.n_points <= 0) return;
        KWH_REQUIRE(s.rows == 1 || s.rows == s.n_points, KWH_EINVAL,
                    std::string(name) + ": signal must have 1 row or one row per source point");
        KWH_REQUIRE(s.len >= 0, KWH_EINVAL, std::string(name) + ": signal length must be >= 0");
        KWH_REQUIRE(s.mode >= 0 && s.mode <= 2, KWH_EINVAL, std::string(name) + ": invalid source mode");
        if (s.delay) {
            KWH_REQUIRE(s.rows == 1, KWH_EINVAL, std::string(name) + ": delayed copies need a single signal row");
            int64_t mx = 0;
            for (int64_t i = 0; i < s.n_points; ++i) {
                KWH_REQUIRE(s.delay[i] >= 0, KWH_EINVAL, std::string(name) + ": delays must be >= 0");
                mx = std::max<int64_t>(mx, s.delay[i]);
            }
            KWH_REQUIRE((int64_t)s.len + mx < (int64_t)1 << 30, KWH_EINVAL, std::string(name) + ": delay too large");
        }
        for (int64_t i = 0; i < s.n_points; ++i) {
            KWH_REQUIRE(s.index[i] >= 0 && s.index[i] < N, KWH_EINVAL, std::string(name) + ": index outside the grid");
            KWH_REQUIRE(i == 0 || s.index[i] > s.index[i - 1], KWH_EINVAL,
                        std::string(name) + ": indices must be ascending and unique");
        }
        const int64_t n = (int64_t)s.rows * s.len;
        for (int64_t i = 0; i < n; ++i)
            KWH_REQUIRE(std::isfinite(s.signal[i]), KWH_EINVAL, std::string(name) + ": signal contains non-finite values");
    };
    src_ok(d->p_source, "p source");
    src_ok(d->u_source[0], "ux source");
    src_ok(d->u_source[1], "uy source");
    src_ok(d->u_source[2], "uz source");
    for (int a = d->dim; a < 3; ++a)
        KWH_REQUIRE(!(d->u_source[a].index && d->u_source[a].signal), KWH_EINVAL,
                    "velocity source component beyond the grid dimension");
    const bool any_src = d->p0.data || (d->p_source.index && d->p_source.signal) ||
                         (d->u_source[0].index && d->u_source[0].signal) ||
                         (d->u_source[1].index && d->u_source[1].signal) ||
                         (d->u_source[2].index && d->u_source[2].signal);
    KWH_REQUIRE(any_src, KWH_EINVAL, "solver requires source.p0 or a time-varying p/u source");
    if (d->sensor_index && d->n_sensor > 0) {
        for (int64_t i = 0; i < d->n_sensor; ++i) {
            KWH_REQUIRE(d->sensor_index[i] >= 0 && d->sensor_index[i] < N, KWH_EINVAL, "sensor index outside the grid");
            KWH_REQUIRE(i == 0 || d->sensor_index[i] > d->sensor_index[i - 1], KWH_EINVAL,
                        "sensor indices must be ascending and unique");
        }
    }
    if (d->sensor_cartesian && d->n_cartesian > 0) {
        for (int64_t p = 0; p < d->n_cartesian; ++p)
            for (int a = 0; a < d->dim; ++a) {
                const double gc = (double)d->sensor_cartesian[a * d->n_cartesian + p] / d->d[a] + (double)(d->n[a] / 2);
                KWH_REQUIRE(gc >= 0 && gc <= d->n[a] - 1, KWH_EINVAL, "Cartesian sensor point outside grid bounds");
            }
    }
}

extern "C" {

int kwh_abi_version(void) { return KWH_ABI_VERSION; }

const char* kwh_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

// Shared with the k-Wave HDF5 I/O translation unit (kwh_h5.cpp); not part of the exported ABI.
__attribute__((visibility("hidden"))) void kwh_set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }
__attribute__((visibility("hidden"))) int kwh_sim_elem_bytes(const kwh_sim* sim) { return sim ? sim->elem_bytes() : 0; }

extern "C" {

// Once per device, with the device current: the address of that device's copy of `scratch`.
static void resolve_scratch(int device) {
    KWH_REQUIRE(device < kMaxDevices, KWH_EUNSUPPORTED, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    KWH_REQUIRE(e == hipSuccess && ptr != nullptr, KWH_EDEVICE,
                std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
}

int kwh_ctx_create(int device_ordinal, kwh_ctx** out) {
    return guarded([&] {
        KWH_REQUIRE(out != nullptr, KWH_EINVAL, "out is NULL");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        KWH_REQUIRE(e == hipSuccess && n > 0, KWH_EDEVICE, "no HIP device available");
        KWH_REQUIRE(device_ordinal >= 0 && device_ordinal < n, KWH_EINVAL, "device ordinal out of range");
        HIPCHK(hipSetDevice(device_ordinal));
        resolve_scratch(device_ordinal);
        *out = new kwh_ctx{device_ordinal};
    });
}

void kwh_ctx_destroy(kwh_ctx* ctx) { delete ctx; }

int kwh_sim_validate(const kwh_sim_desc* desc) {
    return guarded([&] { validate_desc(desc); });
}

int kwh_sim_create(kwh_ctx* ctx, const kwh_sim_desc* desc, kwh_sim** out) {
    return guarded([&] {
        KWH_REQUIRE(ctx && out, KWH_EINVAL, "ctx/out is NULL");
        *out = nullptr;
        validate_desc(desc);
        if (desc->dtype == KWH_FLOAT32) {
            auto s = std::make_unique<SimT<float>>();
            s->setup(*desc, ctx->device);
            *out = s.release();
        } else {
            auto s = std::make_unique<SimT<double>>();
            s->setup(*desc, ctx->device);
            *out = s.release();
        }
    });
}

int kwh_sim_create_axisymmetric(kwh_ctx* ctx, const kwh_sim_desc* desc, kwh_sim** out) {
    return guarded([&] {
        KWH_REQUIRE(ctx && out, KWH_EINVAL, "ctx/out is NULL");
        *out = nullptr;
        validate_desc(desc);
        if (desc->dtype == KWH_FLOAT32) {
            auto s = std::make_unique<SimT<float>>();
            s->axisym = true;
            s->setup(*desc, ctx->device);
            *out = s.release();
        } else {
            auto s = std::make_unique<SimT<double>>();
            s->axisym = true;
            s->setup(*desc, ctx->device);
            *out = s.release();
        }
    });
}

int kwh_sim_run(kwh_sim* sim, int32_t t_begin, int32_t t_end, kwh_step_cb cb, int32_t cb_interval, void* user) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->run(t_begin, t_end, cb, cb_interval, user);
    });
}

int kwh_sim_launch(kwh_sim* sim, int32_t t_begin, int32_t t_end) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->launch(t_begin, t_end);
    });
}

int kwh_sim_sync(kwh_sim* sim) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->sync();
    });
}

int kwh_sim_output_shape(kwh_sim* sim, int32_t field, int64_t shape[3], int32_t* ndim) {
    return guarded([&] {
        KWH_REQUIRE(sim && shape && ndim, KWH_EINVAL, "NULL argument");
        int nd = 0;
        sim->output_shape(field, shape, &nd);
        *ndim = nd;
    });
}

int kwh_sim_read_output(kwh_sim* sim, int32_t field, void* dst, size_t dst_bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim && dst, KWH_EINVAL, "NULL argument");
        sim->read_output(field, dst, dst_bytes);
    });
}

int kwh_sim_read_output_noalloc(kwh_sim* sim, int32_t field, void* dst, size_t dst_bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim && dst, KWH_EINVAL, "NULL argument");
        const int device = sim->device();
        void* ptr = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
        KWH_REQUIRE(ptr != nullptr, KWH_ESTATE, "scratch address of this device not resolved");
        sim->read_output_noalloc(field, dst, dst_bytes, ptr, kScratchBytes);
    });
}

int kwh_sim_steps_done(kwh_sim* sim, int32_t* steps) {
    return guarded([&] {
        KWH_REQUIRE(sim && steps, KWH_EINVAL, "NULL argument");
        *steps = sim->steps_done();
    });
}

int kwh_sim_state_bytes(kwh_sim* sim, size_t* bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim && bytes, KWH_EINVAL, "NULL argument");
        *bytes = sim->state_bytes();
    });
}

int kwh_sim_save_state(kwh_sim* sim, void* dst, size_t bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim && dst, KWH_EINVAL, "NULL argument");
        sim->save_state(dst, bytes);
    });
}

int kwh_sim_load_state(kwh_sim* sim, const void* src, size_t bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim && src, KWH_EINVAL, "NULL argument");
        sim->load_state(src, bytes);
    });
}

int kwh_sim_set_kernel_timing(kwh_sim* sim, int enable) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->set_timing(enable != 0);
    });
}

int kwh_sim_set_kernel_timing_only(kwh_sim* sim, const char* kernel_name) {
    return guarded([&] {
        KWH_REQUIRE(sim && kernel_name, KWH_EINVAL, "NULL argument");
        sim->set_timing_only(kernel_name);
    });
}

int kwh_sim_kernel_stats(kwh_sim* sim, kwh_kernel_stat* stats, int32_t capacity, int32_t* count) {
    return guarded([&] {
        KWH_REQUIRE(sim && count, KWH_EINVAL, "NULL argument");
        *count = sim->kernel_stats(stats, stats ? capacity : 0);
    });
}

int kwh_sim_set_graph(kwh_sim* sim, int enable) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->set_graph(enable != 0);
    });
}

int kwh_sim_set_plane(kwh_sim* sim, int enable) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->set_plane(enable != 0);
    });
}

void kwh_sim_destroy(kwh_sim* sim) { delete sim; }

int kwh_sim_element_traces(kwh_sim* sim, int32_t field, const int32_t* element, const int32_t* delay,
                           int32_t n_elements, double scale, void* dst, size_t dst_bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim && dst, KWH_EINVAL, "NULL argument");
        sim->element_traces(field, element, delay, n_elements, scale, dst, dst_bytes);
    });
}

int kwh_time_reversal(kwh_ctx* ctx, const kwh_sim_desc* desc, const float* recorded, int32_t n_recorded_steps,
                      double compensation_factor, void* p0_out) {
    return guarded([&] {
        KWH_REQUIRE(ctx && desc && recorded && p0_out, KWH_EINVAL, "NULL argument");
        KWH_REQUIRE(desc->sensor_index && desc->n_sensor > 0 && !desc->sensor_cartesian, KWH_EINVAL,
                    "time reversal requires a binary sensor.mask");
        KWH_REQUIRE(n_recorded_steps > 0, KWH_EINVAL, "recordedPressure must be [numSensorPoints, nt]");
        kwh_sim_desc d = *desc;
        d.p0 = kwh_field{nullptr, 0};
        d.p_source = kwh_source_field{desc->sensor_index, desc->n_sensor, recorded, (int32_t)desc->n_sensor,
                                      n_recorded_steps, KWH_SRC_DIRICHLET, nullptr, nullptr};
        for (auto& u : d.u_source) u = kwh_source_field{nullptr, 0, nullptr, 0, 0, 0, nullptr, nullptr};
        d.sensor_index = nullptr;
        d.n_sensor = 0;
        d.sensor_cartesian = nullptr;
        d.n_cartesian = 0;
        d.record = 0;
        d.directivity_angle = nullptr;
        validate_desc(&d);
        std::unique_ptr<kwh_sim> sim;
        if (d.dtype == KWH_FLOAT32) {
            auto s = std::make_unique<SimT<float>>();
            s->setup(d, ctx->device);
            sim = std::move(s);
        } else {
            auto s = std::make_unique<SimT<double>>();
            s->setup(d, ctx->device);
            sim = std::move(s);
        }
        sim->time_reversal_result(compensation_factor, p0_out);
    });
}

int kwh_sim_autotune(kwh_sim* sim, char* report, size_t report_bytes) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        const std::string r = sim->autotune();
        if (report && report_bytes > 0) std::snprintf(report, report_bytes, "%s", r.c_str());
    });
}

int kwh_sim_set_checkpointing(kwh_sim* sim, int32_t every) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->set_checkpointing(every);
    });
}

int kwh_sim_set_nonfinite_check(kwh_sim* sim, int32_t every) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->set_nonfinite_check(every);
    });
}

int kwh_sim_vjp(kwh_sim* sim, const void* dl_dp, int32_t checkpoint_every, void* grad_p0, void* grad_c0) {
    return guarded([&] {
        KWH_REQUIRE(sim, KWH_EINVAL, "sim is NULL");
        sim->vjp(dl_dp, checkpoint_every, grad_p0, grad_c0);
    });
}

int kwh_batch_create(kwh_ctx* ctx, const kwh_sim_desc* descs, int32_t n, kwh_batch** out) {
    return guarded([&] {
        KWH_REQUIRE(ctx && descs && out && n > 0, KWH_EINVAL, "NULL argument or empty batch");
        *out = nullptr;
        for (int i = 0; i < n; ++i) validate_desc(&descs[i]);
        auto b = std::make_unique<kwh_batch>();
        b->dev = ctx->device;
        HIPCHK(hipSetDevice(b->dev));
        for (int i = 0; i < n; ++i) {
            kwh_sim* s = nullptr;
            if (descs[i].dtype == KWH_FLOAT32) {
                auto t = std::make_unique<SimT<float>>();
                t->setup(descs[i], ctx->device);
                s = t.release();
            } else {
                auto t = std::make_unique<SimT<double>>();
                t->setup(descs[i], ctx->device);
                s = t.release();
            }
            b->sims.push_back(s);
        }
        *out = b.release();
    });
}

int kwh_batch_run(kwh_batch* batch, int32_t mode) {
    return guarded([&] {
        KWH_REQUIRE(batch, KWH_EINVAL, "batch is NULL");
        KWH_REQUIRE((mode & ~(KWH_BATCH_GRAPH | KWH_BATCH_FUSED | KWH_BATCH_PLANE)) == 0, KWH_EINVAL, "invalid batch run mode");
        batch->run(mode);
    });
}

int kwh_batch_sim(kwh_batch* batch, int32_t i, kwh_sim** out) {
    return guarded([&] {
        KWH_REQUIRE(batch && out, KWH_EINVAL, "NULL argument");
        KWH_REQUIRE(i >= 0 && i < (int)batch->sims.size(), KWH_EINVAL, "batch index out of range");
        *out = batch->sims[i];
    });
}

int kwh_batch_read_output(kwh_batch* batch, int32_t i, int32_t field, void* dst, size_t dst_bytes) {
    return guarded([&] {
        KWH_REQUIRE(batch && dst, KWH_EINVAL, "NULL argument");
        KWH_REQUIRE(i >= 0 && i < (int)batch->sims.size(), KWH_EINVAL, "batch index out of range");
        batch->sims[i]->read_output(field, dst, dst_bytes);
    });
}

void kwh_batch_destroy(kwh_batch* batch) { delete batch; }

}  // extern "C"
