// Host engine behind include/ptmc.h: run bookkeeping, workgroup planning,
// device buffers, kernel launches.  Every entry point returns a status code;
// the message of the last failure is kept per handle (ptmc_last_error).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../include/ptmc.h"
#include "ptmc_kernels.hpp"
#include "ptmc_nnp_layout.hpp"
#include "ptmc_plugin.hpp"
#include "ptmc_sweep.hpp"

using namespace ptmc;

namespace ptmc {
SweepFn nnp_sweep_fn(int np);
SweepFn nnp_sweep_da_fn();
bool pair_slot_fns(int pk, int kc, bool da, SweepFn *sweep, SweepFn *init);
size_t pair_slot_lds_bytes(bool da);
SweepFn nnp_init_fn(int np);
SweepFn nnp2_sweep_fn();
SweepFn nnp2_init_fn();
SweepFn nnp2_sweep_fn_f32();
SweepFn nnp2_init_fn_f32();
void (*nnp_cycle_fn())(SweepArgs, int);
hipError_t sur_launch_desc(const float *x, const int32_t *n_atoms, float *bd, float r_cut, int n, hipStream_t st);
hipError_t sur_launch_collect(const SweepArgs &a, int n_slots, float *bd, double *be, int32_t *bn, long long base,
                              int cap, hipStream_t st);
hipError_t sur_launch_train(const float *bd, const double *be, const int32_t *bn, const int32_t *idx, const float *w,
                            float *partial, double *pred, int grad, int batch, hipStream_t st);
hipError_t sur_launch_reduce(const float *partial, int batch, float *grad, hipStream_t st);
hipError_t sur_launch_adam(float *w, const float *grad, float *m, float *v, int t, float lr, hipStream_t st);
}  // namespace ptmc

namespace {

// A run's configurations t0 .. t0 + count - 1 held by lane clusters
// local0 .. of one tile (pair potentials)
struct Seg {
    int tile, local0, t0, count;
};

struct RunHost {
    ptmc_run_desc d;
    std::vector<double> beta;
    std::vector<double> x0;  // n_configs * N * 3
    int kl = 1;
    int lanes = 256;  // workgroup width of the run's tile
    RunInfo info{};
    int tile = -1;
    std::vector<Seg> segs;  // one segment, or several in the spread layout
};

struct Launch {
    SweepFn fn = nullptr;    // cycle kernel
    SweepFn init = nullptr;  // initial-energy kernel
    int tile0 = 0, ntiles = 0;
    int run0 = 0, nruns = 0;  // spread layout: the group's runs in the run list
    int wgs = 256;
    int kl = 1;
    size_t lds = 0;
    double ms = 0.0;       // summed duration of this group's timed launches
    long long n_done = 0;  // launches timed
};

// spread layout: one cycle's launches in this many are bracketed by timing
// events (each event pair on a stream costs the GPU about a microsecond
// between the dependent launches, 4 % of a 0.1 ms cycle when every launch
// is timed); the others launch bare
constexpr long long SPREAD_TIME_EVERY = 4;

// start/stop events around one launch, read after the stream is synchronised
struct LaunchTiming {
    int group;
    hipEvent_t a, b;
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void free_()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

struct ptmc_engine {
    ptmc_config cfg{};
    PotParams pot{};
    bool pot_set = false;
    std::vector<RunHost> runs;
    bool prepared = false;
    bool spread = false;  // ladders spread over workgroups (cfg.workgroup_lanes > 0)
    int64_t recompute_applied = -1;  // spread layout: sweep whose pending recompute a launch applied
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // side streams (default priority) that run launch groups concurrently
    std::vector<hipStream_t> side;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    double kernel_ms = 0.0;
    long long n_launch = 0;
    long long spread_launch = 0;  // spread-layout cycle launches so far (timing sample)
    std::vector<hipEvent_t> ev_pool;  // every timing event created (destroyed with the handle)
    std::vector<hipEvent_t> ev_free;  // those not attached to a pending launch
    std::vector<LaunchTiming> pending;

    // plan
    std::vector<TileInfo> tiles;
    std::vector<Launch> launches;
    int pk = PK_EVEN_H, kc = 8;
    int n_slots = 0, n_temps = 0;
    int64_t n_coord = 0;
    int64_t n_comp = 0;  // EAM phi/rho elements
    int hist_stride = 0;

    // device state
    DevBuf<TileInfo> d_tiles;
    DevBuf<RunInfo> d_runs;
    DevBuf<int32_t> d_run_list;  // spread layout: runs of each launch group, group after group
    std::vector<int32_t> run_list;
    DevBuf<int32_t> d_slot_run, d_slot_temp, d_run_error;
    DevBuf<double> d_slot_en, d_slot_box, d_slot_boxh, d_slot_drift;
    DevBuf<unsigned char> d_coords, d_comps;
    DevBuf<double> d_beta, d_displ, d_maxbox, d_ham;
    DevBuf<long long> d_count_atom, d_count_vol, d_count_exc;
    // boxes with a height: per temperature max height, length/height bound,
    // separated-move steps (xy, z) and acceptance counters
    DevBuf<double> d_maxboxh, d_lh, d_displ_sep;
    // LookupTablePotential table, FP64 and binary32 copies
    std::vector<double> lut_host;
    DevBuf<double> d_lut64;
    DevBuf<float> d_lut32;
    DevBuf<long long> d_count_volxy, d_count_volz;
    bool separated = false;
    DevBuf<uint32_t> d_hist;
    DevBuf<double> d_trace_de;
    DevBuf<int8_t> d_trace_acc;
    DevBuf<int32_t> d_trace_idx;
    DevBuf<double> d_trace_aux;
    int trace_cycles = 0;
    int trace_pending = 0;
    int trace_nsteps = 0;

    // neural-network potential (one workgroup per slot)
    struct {
        std::vector<float> rad, geta, afn, rowoff, w;
        std::vector<int32_t> grp, abuild, gdesc, wgroups;
        int n_rad = 0, n_ang = 0, nf = 0, h1 = 0, h2 = 0, n_rad_active = 0, n_groups = 0, q10 = 0;
        double r_cut = 0.0, hardcore2 = 0.0;
        float b3 = 0.0f;
        // FP64 tables as given, for the plugin functions
        std::vector<double> rad64, ang64, w64;
    } nnp;
    DevBuf<float> d_nnp_rad, d_nnp_geta, d_nnp_afn, d_nnp_rowoff, d_nnp_w, d_feat, d_eatom;
    DevBuf<int32_t> d_nnp_grp, d_nnp_abuild, d_nnp_gdesc, d_nnp_wgroups, d_temp_slot;
    // plugin functions: FP64 network tables (uploaded on first use) and a
    // grow-only arena for each call's inputs, outputs and workspace
    DevBuf<double> d_plug_rad, d_plug_ang, d_plug_w, d_plug;
    bool plug_nnp_up = false;
    NnpTables nnp_dev{};
    size_t nnp_lds = 0;
    int nnp_np = NNP_NP;  // atoms per slot of the network-potential engine (64 or 128)

    // two-species network potential, NNVT (one workgroup per slot, FP64)
    bool nnvt = false;
    struct {
        std::vector<double> rad, ang, p1, p2;
        std::vector<int32_t> grp, gfn;
        int n1 = 0, n2 = 0, n_swaps = 0, n_rad = 0, n_ang = 0, ng = 0, h1 = 0, h2 = 0, n_groups = 0;
        int act[4] = {0, 0, 0, 0};
        double r_cut = 0.0;
    } nnp2;
    DevBuf<double> d_n2_rad, d_n2_ang, d_n2_p1, d_n2_p2, d_n2_feat, d_n2_eatom;
    DevBuf<int32_t> d_n2_grp, d_n2_gfn;
    Nnp2Tables nnp2_dev{};
    size_t nnp2_lds = 0;

    // learned surrogate / delayed acceptance
    struct {
        bool init = false, da = false;
        int cap = 0, collect_every = 0, t = 0;
        long long n_samples = 0;
        uint64_t seed = 0;
    } sur;
    DevBuf<float> d_sur_w, d_sur_m, d_sur_v, d_sur_grad, d_sur_partial, d_sur_D, d_sur_x;
    DevBuf<double> d_sur_E, d_sur_pred;
    DevBuf<int32_t> d_sur_N, d_sur_idx;
    DevBuf<long long> d_da_stats;

    int64_t sweep_counter = 0;  // global cycle counter (Philox word 1)
    int64_t phase_counter = 0;  // phase-local cycle index
    int production = 0;
};

namespace {

int fail(ptmc_engine *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(h, PTMC_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));     \
    } while (0)

template <typename T>
int dalloc(ptmc_engine *h, DevBuf<T> &b, size_t n)
{
    b.free_();
    if (n == 0) n = 1;
    HIPCHK(h, hipMalloc((void **)&b.p, n * sizeof(T)));
    b.n = n;
    HIPCHK(h, hipMemsetAsync(b.p, 0, n * sizeof(T), h->stream));
    return PTMC_OK;
}

template <typename T>
int upload(ptmc_engine *h, DevBuf<T> &b, const std::vector<T> &v)
{
    int rc = dalloc(h, b, v.size());
    if (rc) return rc;
    if (!v.empty()) HIPCHK(h, hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
    return PTMC_OK;
}

int default_lanes(int n_atoms, int precision)
{
    // lanes per configuration: keep ~13-16 atoms per lane in binary32,
    // ~8-13 in FP64 (twice the LDS per atom)
    int per = precision == 1 ? 16 : 13;
    int kl = 1;
    while (kl < 16 && (n_atoms + kl - 1) / kl > per) kl *= 2;
    return kl;
}

// bc is the configured boundary: rectangular boxes run the cubic kernels
KernelPair lookup(int precision, int bc, int ens, int pot, int pk, int kc, int kl, int lanes)
{
    if (pot == POT_EAM) {
        if (bc != BC_SPHERE && bc != BC_CUBIC) return {};
        return ens == ENS_NVT ? lookup_eam(precision, bc, kl, lanes) : KernelPair{};
    }
    if (bc == BC_RECT) bc = BC_CUBIC;
    if (precision == 0) {
        if (bc == BC_SPHERE && ens == ENS_NVT) return lookup_f64_sphere_nvt(pk, kc, kl, lanes);
        if (bc == BC_CUBIC && ens == ENS_NVT) return lookup_f64_cubic_nvt(pk, kc, kl, lanes);
        if (bc == BC_CUBIC && ens == ENS_NPT) return lookup_f64_cubic_npt(pk, kc, kl, lanes);
        if (bc == BC_RHOMB && ens == ENS_NVT) return lookup_f64_rhomb_nvt(pk, kc, kl, lanes);
        if (bc == BC_RHOMB && ens == ENS_NPT) return lookup_f64_rhomb_npt(pk, kc, kl, lanes);
    } else {
        if (bc == BC_SPHERE && ens == ENS_NVT) return lookup_f32_sphere_nvt(pk, kc, kl, lanes);
        if (bc == BC_CUBIC && ens == ENS_NVT) return lookup_f32_cubic_nvt(pk, kc, kl, lanes);
        if (bc == BC_CUBIC && ens == ENS_NPT) return lookup_f32_cubic_npt(pk, kc, kl, lanes);
        if (bc == BC_RHOMB && ens == ENS_NVT) return lookup_f32_rhomb_nvt(pk, kc, kl, lanes);
        if (bc == BC_RHOMB && ens == ENS_NPT) return lookup_f32_rhomb_npt(pk, kc, kl, lanes);
    }
    return {};
}

// LDS of one tile: positions [3][A][lanes] (EAM adds two phi/rho buffers
// [2][A][2][lanes]) and the per-slot exchange scratch.
size_t tile_lds(size_t rsize, int n_atoms, int kl, int lanes, bool eam)
{
    const int A = (n_atoms + kl - 1) / kl;
    // positions; per slot: exchange energy, volume, label, slot map; FP64 pair
    // sweeps also 64 B of per-temperature state (four counters, two sums,
    // step size, beta)
    const bool rs = !eam && rsize == sizeof(double);
    return rsize * (eam ? 7 : 3) * A * lanes + (size_t)(lanes / kl) * (8 + 8 + 4 + 4 + (rs ? 64 : 0));
}

int check_run(ptmc_engine *h, int32_t run)
{
    if (!h) return PTMC_E_ARG;
    if (run < 0 || run >= (int)h->runs.size()) return fail(h, PTMC_E_ARG, "run index %d out of range", run);
    if (!h->prepared) return fail(h, PTMC_E_STATE, "engine not prepared");
    return PTMC_OK;
}

// Select pair evaluation and coefficient padding (trailing zero
// coefficients leave Horner and the reference loop bit-identical).
int choose_pair(ptmc_engine *h)
{
    const PotParams &p = h->pot;
    const int K = p.ncoeff;
    if (p.kind == POT_EAM) {
        if (h->cfg.ensemble != ENS_NVT) return fail(h, PTMC_E_UNSUPPORTED, "EAM runs in the NVT ensemble only");
        if (h->cfg.boundary != BC_SPHERE && h->cfg.boundary != BC_CUBIC)
            return fail(h, PTMC_E_UNSUPPORTED, "EAM runs in a spherical container or a cubic box");
        if (h->cfg.pair_form == 1) return fail(h, PTMC_E_UNSUPPORTED, "EAM has no reference pair form");
        return PTMC_OK;
    }
    if (p.kind == POT_LUT) {
        h->pk = PK_LUT;
        h->kc = 1;
        std::vector<float> f(h->lut_host.begin(), h->lut_host.end());
        int rc = upload(h, h->d_lut64, h->lut_host);
        if (rc) return rc;
        if ((rc = upload(h, h->d_lut32, f))) return rc;
        h->pot.lut64 = h->d_lut64.p;
        h->pot.lut32 = h->d_lut32.p;
        return PTMC_OK;
    }
    if (p.kind == POT_ELJB) {
        // one form: FP64 follows the reference's steps, binary32 its own
        if (K > 8) return fail(h, PTMC_E_UNSUPPORTED, "at most 8 ELJPotentialB coefficients");
        h->pk = PK_ELJB;
        h->kc = 8;
        return PTMC_OK;
    }
    if (h->cfg.pair_form == 1) {
        if (h->cfg.precision != 0) return fail(h, PTMC_E_UNSUPPORTED, "reference pair form requires FP64");
        if (K > 8) return fail(h, PTMC_E_UNSUPPORTED, "reference pair form supports up to 8 coefficients");
        h->pk = p.kind == POT_ELJ_EVEN ? PK_EVEN_R : PK_ODD_R;
        h->kc = 8;
        return PTMC_OK;
    }
    if (p.kind == POT_ELJ) {
        if (K > 8) return fail(h, PTMC_E_UNSUPPORTED, "at most 8 ELJ coefficients");
        h->pk = PK_ODD_H;
        h->kc = 8;
        return PTMC_OK;
    }
    if (K > 8) return fail(h, PTMC_E_UNSUPPORTED, "at most 8 ELJ coefficients");
    if (h->cfg.precision == 1 && K == 4 && p.coeff[1] == 0.0 && p.coeff[2] == 0.0) {
        h->pk = PK_LJ;  // Lennard-Jones written as ELJPotentialEven{4}([c6, 0, 0, c12])
        h->kc = 4;
        return PTMC_OK;
    }
    h->pk = PK_EVEN_H;
    h->kc = K <= 6 ? 6 : 8;
    return PTMC_OK;
}

// Spread layout (cfg.workgroup_lanes = W): runs of equal (lanes per
// configuration, N) share tiles of W lanes; a ladder longer than a tile is cut
// into nearly equal parts on consecutive tiles, so a run's slots stay
// contiguous.  The runs are dealt into cfg.spread_streams groups of nearly
// equal work (largest first, each to the lightest group); a group's tiles go
// into one launch of the spread kernel, ordered by decreasing work so that
// the heaviest workgroups are dispatched first, and the groups run on their
// own streams.
int plan_spread(ptmc_engine *h)
{
    const int W = h->cfg.workgroup_lanes;
    const int prec = h->cfg.precision;
    const size_t rsize = prec == 0 ? sizeof(double) : sizeof(float);
    if (W != 64 && W != 128 && W != 256)
        return fail(h, PTMC_E_ARG, "workgroup_lanes must be 0, 64, 128 or 256");
    const bool sphere = h->cfg.boundary == BC_SPHERE;
    const bool box_ok = h->cfg.boundary == BC_CUBIC || h->cfg.boundary == BC_RECT;
    if (!((sphere && h->cfg.ensemble == ENS_NVT) || (box_ok && !h->separated)) || h->nnvt ||
        h->pot.kind == POT_EAM || h->pot.kind == POT_ELJB || h->pot.kind == POT_LUT || h->cfg.pair_form != 0)
        return fail(h, PTMC_E_UNSUPPORTED,
                    "the spread layout covers Horner-form pair potentials in a spherical container (NVT) or a "
                    "cubic / rectangular box (NVT, NPT with uniform volume moves)");
    if (h->cfg.spread_streams < 0 || h->cfg.spread_streams > 4)
        return fail(h, PTMC_E_ARG, "spread_streams must be 0 to 4");
    const KernelPair kp = lookup_spread(prec, h->pk, h->kc, W, sphere ? BC_SPHERE : BC_CUBIC, h->cfg.ensemble,
                                        h->cfg.exchange_every > 1);
    if (!kp.sweep) return fail(h, PTMC_E_UNSUPPORTED, "no spread kernel for pair form %d/%d", h->pk, h->kc);
    int dev = 0;
    HIPCHK(h, hipGetDevice(&dev));
    int max_lds = 0;
    HIPCHK(h, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));

    const int n_runs = (int)h->runs.size();
    for (int r = 0; r < n_runs; ++r) {
        RunHost &rh = h->runs[r];
        const int N = rh.d.n_atoms;
        const int kl = rh.d.lanes_per_replica > 0 ? rh.d.lanes_per_replica
                       : h->cfg.lanes_per_replica > 0 ? h->cfg.lanes_per_replica
                                                      : default_lanes(N, prec);
        if (kl < 1 || kl > 64 || (kl & (kl - 1)) != 0 || kl > W)
            return fail(h, PTMC_E_ARG, "run %d: lanes per configuration must be a power of two up to 64 (at most %d)",
                        r, W);
        if (rh.d.n_temps > SPREAD_MAX_TEMPS)
            return fail(h, PTMC_E_UNSUPPORTED, "run %d: at most %d temperatures in the spread layout", r,
                        SPREAD_MAX_TEMPS);
        const size_t lds = rsize * 3 * (size_t)((N + kl - 1) / kl) * W;
        if (lds > (size_t)max_lds)
            return fail(h, PTMC_E_UNSUPPORTED, "run %d: a tile needs %zu B of LDS (device limit %d)", r, lds, max_lds);
        rh.kl = kl;
        rh.lanes = W;
        rh.segs.clear();
    }
    // runs into groups of nearly equal work (configurations x moves x slabs)
    const int G = std::max(1, std::min(h->cfg.spread_streams, n_runs));
    auto run_work = [&](int r) {
        const RunHost &rh = h->runs[r];
        return (double)rh.d.n_temps * rh.d.n_atoms * ((rh.d.n_atoms + rh.kl - 1) / rh.kl + 3) * rh.kl;
    };
    std::vector<int> by_work(n_runs);
    for (int r = 0; r < n_runs; ++r) by_work[r] = r;
    std::stable_sort(by_work.begin(), by_work.end(), [&](int x, int y) { return run_work(x) > run_work(y); });
    std::vector<std::vector<int>> grp(G);
    std::vector<double> load(G, 0.0);
    for (int r : by_work) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        grp[g].push_back(r);
        load[g] += run_work(r);
    }
    for (auto &v : grp) std::sort(v.begin(), v.end());

    struct Part {
        int run, t0, count;
    };
    h->tiles.clear();
    h->launches.clear();
    h->run_list.clear();
    std::vector<TileInfo> ordered;
    int slot = 0;
    int64_t coord = 0;
    size_t maxlds = 0;
    for (int g = 0; g < G; ++g) {
        std::map<std::pair<int, int>, std::vector<int>> groups;  // (kl, N) -> runs
        for (int r : grp[g]) groups[{h->runs[r].kl, h->runs[r].d.n_atoms}].push_back(r);
        std::vector<std::vector<Part>> packs;
        std::vector<int> pack_kl, pack_n;
        for (auto &gg : groups) {
            const int kl = gg.first.first, N = gg.first.second, cap = W / kl;
            std::vector<int> rs = gg.second;
            std::stable_sort(rs.begin(), rs.end(),
                             [&](int a, int b) { return h->runs[a].d.n_temps > h->runs[b].d.n_temps; });
            const size_t first = packs.size();
            std::vector<int> used;
            for (int r : rs) {
                const int R = h->runs[r].d.n_temps;
                if (R > cap) {
                    // consecutive tiles of its own, parts of nearly equal size
                    const int k = (R + cap - 1) / cap;
                    int t0 = 0;
                    for (int p = 0; p < k; ++p) {
                        const int c = R / k + (p < R % k ? 1 : 0);
                        packs.push_back({Part{r, t0, c}});
                        pack_kl.push_back(kl);
                        pack_n.push_back(N);
                        used.push_back(cap);
                        t0 += c;
                    }
                    continue;
                }
                int t = -1;
                for (size_t k = first; k < packs.size(); ++k)
                    if (used[k - first] + R <= cap) {
                        t = (int)k;
                        break;
                    }
                if (t < 0) {
                    packs.push_back({});
                    pack_kl.push_back(kl);
                    pack_n.push_back(N);
                    used.push_back(0);
                    t = (int)packs.size() - 1;
                }
                packs[t].push_back(Part{r, 0, R});
                used[t - first] += R;
            }
        }
        // slots in pack order: a run's parts sit on consecutive tiles
        std::vector<TileInfo> tiles;
        std::vector<std::vector<std::pair<int, Seg>>> segs_of(packs.size());
        for (size_t k = 0; k < packs.size(); ++k) {
            TileInfo ti{};
            ti.slot0 = slot;
            ti.n_atoms = pack_n[k];
            ti.kl = pack_kl[k];
            ti.n_slab = (ti.n_atoms + ti.kl - 1) / ti.kl;
            int local = 0;
            for (const Part &p : packs[k]) {
                RunHost &rh = h->runs[p.run];
                if (p.t0 == 0) {
                    rh.info.slot_base = slot + local;
                    rh.info.local_base = local;
                }
                segs_of[k].push_back({p.run, Seg{-1, local, p.t0, p.count}});
                local += p.count;
            }
            ti.n_slots = local;
            slot += local;
            tiles.push_back(ti);
        }
        // heaviest tiles first: configurations x moves x (slabs + per-move overhead)
        std::vector<int> order(tiles.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
        auto work = [&](int k) { return (long long)tiles[k].n_slots * tiles[k].n_atoms * (tiles[k].n_slab + 3); };
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
        Launch L;
        L.tile0 = (int)ordered.size();
        L.ntiles = (int)tiles.size();
        L.wgs = W;
        L.kl = 0;
        L.fn = kp.sweep;
        L.init = kp.init;
        L.run0 = (int)h->run_list.size();
        L.nruns = (int)grp[g].size();
        for (int k : order) {
            TileInfo ti = tiles[k];
            ti.coord_off = coord;
            coord += (int64_t)3 * ti.n_slab * W;
            L.lds = std::max(L.lds, rsize * 3 * (size_t)ti.n_slab * W);
            for (auto &rs : segs_of[k]) {
                Seg sg = rs.second;
                sg.tile = (int)ordered.size();
                RunHost &rh = h->runs[rs.first];
                if (sg.t0 == 0) rh.tile = sg.tile;
                rh.segs.push_back(sg);
            }
            ordered.push_back(ti);
        }
        for (int r : grp[g]) h->run_list.push_back(r);
        maxlds = std::max(maxlds, L.lds);
        h->launches.push_back(L);
    }
    for (auto &rh : h->runs)
        std::sort(rh.segs.begin(), rh.segs.end(), [](const Seg &x, const Seg &y) { return x.t0 < y.t0; });
    h->tiles = ordered;
    h->n_slots = slot;
    h->n_coord = coord;
    h->n_comp = 0;
    HIPCHK(h, hipFuncSetAttribute((const void *)kp.sweep, hipFuncAttributeMaxDynamicSharedMemorySize, (int)maxlds));
    HIPCHK(h, hipFuncSetAttribute((const void *)kp.init, hipFuncAttributeMaxDynamicSharedMemorySize, (int)maxlds));
    int rc = upload(h, h->d_run_list, h->run_list);
    return rc;
}

int plan(ptmc_engine *h)
{
    if (h->cfg.workgroup_lanes > 0) {
        h->spread = true;
        return plan_spread(h);
    }
    const int prec = h->cfg.precision;
    const size_t rsize = prec == 0 ? sizeof(double) : sizeof(float);
    const bool eam = h->pot.kind == POT_EAM;
    int dev = 0;
    HIPCHK(h, hipGetDevice(&dev));
    int max_lds = 0;
    HIPCHK(h, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));

    // group runs by (lanes per configuration, workgroup width, N)
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    for (int r = 0; r < (int)h->runs.size(); ++r) {
        RunHost &rh = h->runs[r];
        const int N = rh.d.n_atoms, R = rh.d.n_temps;
        const int req = rh.d.lanes_per_replica > 0 ? rh.d.lanes_per_replica
                        : h->cfg.lanes_per_replica > 0 ? h->cfg.lanes_per_replica
                                                       : default_lanes(N, prec);
        if (req > 16)
            return fail(h, PTMC_E_UNSUPPORTED, "run %d: %d lanes per configuration need the spread layout "
                        "(workgroup_lanes > 0)", r, req);
        // smallest built workgroup width that holds the whole ladder; more
        // lanes per configuration when the slab does not fit LDS
        int kl = -1, lanes = 0;
        for (int w : {LANES_SMALL, LANES_HALF, LANES_MID, LANES_LARGE}) {
            if (eam && w != LANES_SMALL) continue;
            for (int k = (w != LANES_SMALL && req < 2) ? 2 : req; k <= 16 && R * k <= w; k *= 2) {
                if (tile_lds(rsize, N, k, w, eam) <= (size_t)max_lds &&
                    lookup(prec, h->cfg.boundary, h->cfg.ensemble, h->pot.kind, h->pk, h->kc, k, w).sweep) {
                    kl = k;
                    lanes = w;
                    break;
                }
            }
            if (kl > 0) break;
        }
        if (kl < 0)
            return fail(h, PTMC_E_UNSUPPORTED, "run %d: %d temperatures x %d lanes per configuration do not fit one "
                        "workgroup", r, R, req);
        rh.kl = kl;
        rh.lanes = lanes;
        groups[{kl, lanes, N}].push_back(r);
    }

    h->tiles.clear();
    h->launches.clear();
    int slot = 0;
    int64_t coord = 0, comp = 0;
    // one launch per (lanes per configuration, workgroup width)
    std::map<std::pair<int, int>, std::vector<int>> tiles_by_kl;
    for (auto &g : groups) {
        const int kl = std::get<0>(g.first), wgs = std::get<1>(g.first), N = std::get<2>(g.first);
        std::vector<int> &rs = g.second;
        std::stable_sort(rs.begin(), rs.end(),
                         [&](int a, int b) { return h->runs[a].d.n_temps > h->runs[b].d.n_temps; });
        const int cap = wgs / kl;
        // first-fit of whole ladders into tiles of cap slots
        std::vector<std::vector<int>> packs;
        std::vector<int> used;
        for (int r : rs) {
            const int R = h->runs[r].d.n_temps;
            int t = -1;
            for (size_t k = 0; k < packs.size(); ++k)
                if (used[k] + R <= cap) { t = (int)k; break; }
            if (t < 0) {
                packs.push_back({});
                used.push_back(0);
                t = (int)packs.size() - 1;
            }
            packs[t].push_back(r);
            used[t] += R;
        }
        for (size_t k = 0; k < packs.size(); ++k) {
            TileInfo ti{};
            ti.slot0 = slot;
            ti.n_atoms = N;
            ti.n_slab = (N + kl - 1) / kl;
            ti.kl = kl;
            ti.coord_off = 0;  // filled below (needs the launch's wgs)
            int local = 0;
            for (int r : packs[k]) {
                RunHost &rh = h->runs[r];
                rh.info.slot_base = slot + local;
                rh.info.local_base = local;
                rh.tile = (int)h->tiles.size();
                rh.segs = {Seg{rh.tile, local, 0, rh.d.n_temps}};
                local += rh.d.n_temps;
            }
            ti.n_slots = local;
            slot += local;
            tiles_by_kl[{kl, wgs}].push_back((int)h->tiles.size());
            h->tiles.push_back(ti);
        }
    }
    h->n_slots = slot;

    // reorder tiles so each launch is contiguous
    std::vector<TileInfo> ordered;
    std::vector<int> old_to_new(h->tiles.size());
    for (auto &kv : tiles_by_kl) {
        const int kl = kv.first.first;
        Launch L;
        L.tile0 = (int)ordered.size();
        L.ntiles = (int)kv.second.size();
        L.wgs = kv.first.second;
        L.kl = kl;
        const KernelPair kp = lookup(prec, h->cfg.boundary, h->cfg.ensemble, h->pot.kind, h->pk, h->kc, kl, L.wgs);
        L.fn = kp.sweep;
        L.init = kp.init;
        if (!L.fn || !L.init)
            return fail(h, PTMC_E_UNSUPPORTED, "no kernel for precision %d boundary %d ensemble %d pair %d/%d lanes %d",
                        prec, h->cfg.boundary, h->cfg.ensemble, h->pk, h->kc, kl);
        size_t maxlds = 0;
        for (int t : kv.second) {
            TileInfo ti = h->tiles[t];
            ti.coord_off = coord;
            coord += (int64_t)3 * ti.n_slab * L.wgs;
            ti.comp_off = comp;
            if (eam) comp += (int64_t)2 * ti.n_slab * L.wgs;
            size_t lds = tile_lds(rsize, ti.n_atoms, kl, L.wgs, eam);
            maxlds = std::max(maxlds, lds);
            old_to_new[t] = (int)ordered.size();
            ordered.push_back(ti);
        }
        if ((long long)maxlds > max_lds)
            return fail(h, PTMC_E_UNSUPPORTED, "tile needs %zu B of LDS (device limit %d); use more lanes per configuration",
                        maxlds, max_lds);
        L.lds = maxlds;
        HIPCHK(h, hipFuncSetAttribute((const void *)L.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)maxlds));
        HIPCHK(h, hipFuncSetAttribute((const void *)L.init, hipFuncAttributeMaxDynamicSharedMemorySize, (int)maxlds));
        h->launches.push_back(L);
    }
    // heaviest launch group first (per-cycle work of its largest tile ~ N * slab)
    std::stable_sort(h->launches.begin(), h->launches.end(), [&](const Launch &x, const Launch &y) {
        auto work = [&](const Launch &L) {
            long long w = 0;
            for (int t = L.tile0; t < L.tile0 + L.ntiles; ++t)
                w = std::max<long long>(w, (long long)ordered[t].n_atoms * ordered[t].n_slab);
            return w;
        };
        return work(x) > work(y);
    });
    for (auto &rh : h->runs) {
        rh.tile = old_to_new[rh.tile];
        for (auto &g : rh.segs) g.tile = old_to_new[g.tile];
    }
    h->tiles = ordered;
    h->n_coord = coord;
    h->n_comp = comp;
    // slot0 of reordered tiles keeps the values assigned above (slots stay with
    // their runs); coordinates are addressed by tile.coord_off.
    return PTMC_OK;
}

template <typename Real>
void scatter_coords(ptmc_engine *h, std::vector<Real> &buf)
{
    buf.assign((size_t)h->n_coord, (Real)0);
    const Real far = h->cfg.boundary == BC_SPHERE ? (Real)(sizeof(Real) == 4 ? 1.0e18 : 1.0e100) : (Real)0;
    for (auto &rh : h->runs)
        for (const Seg &g : rh.segs) {
            const TileInfo &ti = h->tiles[g.tile];
            const int N = rh.d.n_atoms, A = ti.n_slab, kl = rh.kl;
            for (int t = g.t0; t < g.t0 + g.count; ++t) {
                const int lg = g.local0 + (t - g.t0);  // slot t holds temperature t at start
                const double *x = &rh.x0[(size_t)(t % rh.d.n_configs) * N * 3];
                for (int j = 0; j < A * kl; ++j) {
                    const int s = j / kl, lane = lg * kl + j % kl;
                    for (int c = 0; c < 3; ++c)
                        buf[(size_t)ti.coord_off + ((size_t)s * 3 + c) * rh.lanes + lane] =
                            j < N ? (Real)x[3 * j + c] : far;
                }
            }
        }
}

int launch_wgs(ptmc_engine *h, int tile)
{
    for (auto &L : h->launches)
        if (tile >= L.tile0 && tile < L.tile0 + L.ntiles) return L.wgs;
    return 256;
}

bool is_pair_slot(const ptmc_engine *h);

// Cutoff of the surrogate's radial descriptors: the network potential's own;
// for pair potentials 6 A (11.338 bohr), the descriptors' grid starting at
// 2 bohr (atomic units, as the neon and copper set-ups of the reference).
double sur_cutoff_of(const ptmc_engine *h) { return is_pair_slot(h) ? 11.338 : h->nnp.r_cut; }

SweepArgs base_args(ptmc_engine *h)
{
    SweepArgs a{};
    a.tiles = h->d_tiles.p;
    a.runs = h->d_runs.p;
    a.slot_run = h->d_slot_run.p;
    a.slot_temp = h->d_slot_temp.p;
    a.slot_en = h->d_slot_en.p;
    a.slot_box = h->d_slot_box.p;
    a.slot_boxh = h->d_slot_boxh.p;
    a.slot_drift = h->d_slot_drift.p;
    a.coords = h->d_coords.p;
    a.comps = h->d_comps.p;
    a.t_beta = h->d_beta.p;
    a.t_displ = h->d_displ.p;
    a.t_maxbox = h->d_maxbox.p;
    a.t_maxboxh = h->d_maxboxh.p;
    a.t_lh = h->d_lh.p;
    a.t_displ_sep = h->d_displ_sep.p;
    a.t_count_volxy = h->d_count_volxy.p;
    a.t_count_volz = h->d_count_volz.p;
    a.box_kind = h->cfg.boundary == BC_RECT ? BC_RECT : BC_CUBIC;
    a.separated = h->separated ? 1 : 0;
    a.t_count_atom = h->d_count_atom.p;
    a.t_count_vol = h->d_count_vol.p;
    a.t_count_exc = h->d_count_exc.p;
    a.t_ham = h->d_ham.p;
    a.t_hist = h->d_hist.p;
    a.run_error = h->d_run_error.p;
    a.pot = h->pot;
    a.seed = h->cfg.seed;
    a.ensemble = h->cfg.ensemble;
    a.exchange_mode = h->cfg.exchange_mode;
    a.exchange_every = h->cfg.exchange_every > 1 ? h->cfg.exchange_every : 1;
    a.n_adjust = h->cfg.n_adjust;
    a.mc_sample = h->cfg.mc_sample;
    a.n_bin = h->cfg.n_bin;
    a.hist_stride = h->hist_stride;
    a.recompute_every = h->cfg.recompute_every;
    a.min_acc = h->cfg.min_acc;
    a.max_acc = h->cfg.max_acc;
    a.nnp = h->nnp_dev;
    a.nnp2 = h->nnp2_dev;
    a.feat = h->d_feat.p;
    a.eatom = h->d_eatom.p;
    a.temp_slot = h->d_temp_slot.p;
    a.sur_w = h->d_sur_w.p;
    a.da = h->sur.da ? 1 : 0;
    a.sur_rcut = (float)sur_cutoff_of(h);
    a.da_stats = h->d_da_stats.p;
    return a;
}

bool is_nnp(const ptmc_engine *h) { return h->pot.kind == POT_NNP; }
bool is_nnp2(const ptmc_engine *h) { return h->pot.kind == POT_NNP2; }
// potentials run one workgroup per slot (label exchange between cycles)
// pair potentials one workgroup per configuration (cfg.pair_slots)
bool is_pair_slot(const ptmc_engine *h) { return h->cfg.pair_slots != 0 && !is_nnp(h) && !is_nnp2(h); }
bool is_slot(const ptmc_engine *h) { return is_nnp(h) || is_nnp2(h) || is_pair_slot(h); }
// atom stride of a slot engine's positions, features and atom energies
int slot_np(const ptmc_engine *h) { return is_nnp(h) ? h->nnp_np : NNP_NP; }
// slot engines keeping binary32 positions [slot][3][slot_np]
bool slot_f32(const ptmc_engine *h)
{
    return is_nnp(h) || is_pair_slot(h) || (is_nnp2(h) && h->cfg.precision == 1);
}

// Pair slots: one workgroup per configuration, positions [slot][3][NNP_NP]
// in binary32, the surrogate's layout for delayed acceptance.
int prepare_pair_slot(ptmc_engine *h)
{
    if (h->cfg.precision != 1) return fail(h, PTMC_E_UNSUPPORTED, "pair slots run in binary32 (precision 1)");
    if (h->cfg.ensemble != ENS_NVT || h->cfg.boundary != BC_SPHERE)
        return fail(h, PTMC_E_UNSUPPORTED, "pair slots run NVT in a spherical container");
    SweepFn fs = nullptr, fi = nullptr;
    if (h->pot.kind == POT_EAM || !pair_slot_fns(h->pk, h->kc, false, &fs, &fi))
        return fail(h, PTMC_E_UNSUPPORTED, "pair slots take the Lennard-Jones or even extended-LJ (up to 8 "
                    "coefficients) pair forms");
    h->nnp_lds = pair_slot_lds_bytes(false);
    int slot = 0;
    for (auto &rh : h->runs) {
        if (rh.d.n_atoms > NNP_NP) return fail(h, PTMC_E_UNSUPPORTED, "at most %d atoms in pair slots", NNP_NP);
        rh.info.slot_base = slot;
        rh.info.local_base = 0;
        rh.kl = PSL_NT;
        rh.lanes = PSL_NT;
        rh.tile = -1;
        slot += rh.d.n_temps;
    }
    h->n_slots = slot;
    h->tiles.clear();
    h->launches.clear();
    Launch L;
    L.fn = fs;
    L.init = fi;
    L.ntiles = slot;
    L.wgs = PSL_NT;
    L.kl = PSL_NT;
    L.lds = h->nnp_lds;
    h->launches.push_back(L);
    h->n_coord = (int64_t)slot * 3 * NNP_NP;
    h->n_comp = 0;
    return PTMC_OK;
}

// Slots of the two-species engine: as prepare_nnp, FP64 positions
// [slot][3][NNP_NP], features [slot][NNP_NP][ng] and atom energies in HBM.
int prepare_nnp2(ptmc_engine *h)
{
    auto &n = h->nnp2;
    if (h->cfg.boundary != BC_SPHERE) return fail(h, PTMC_E_UNSUPPORTED, "the two-species network potential runs in a spherical container");
    int dev = 0, max_lds = 0;
    HIPCHK(h, hipGetDevice(&dev));
    HIPCHK(h, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int N = n.n1 + n.n2;
    const bool f32 = h->cfg.precision == 1;
    h->nnp2_lds = nnp2_lds_bytes(N, n.ng, n.n_rad + n.n_ang, n.h1, n.h2, f32 ? 4 : 8);
    if ((long long)h->nnp2_lds > max_lds)
        return fail(h, PTMC_E_UNSUPPORTED, "%d atoms with %d feature rows need %zu B of LDS (%d available)", N, n.ng,
                    h->nnp2_lds, max_lds);
    const void *fs = (const void *)(f32 ? nnp2_sweep_fn_f32() : nnp2_sweep_fn());
    const void *fi = (const void *)(f32 ? nnp2_init_fn_f32() : nnp2_init_fn());
    HIPCHK(h, hipFuncSetAttribute(fs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->nnp2_lds));
    HIPCHK(h, hipFuncSetAttribute(fi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->nnp2_lds));
    int slot = 0;
    for (auto &rh : h->runs) {
        if (rh.d.n_atoms != N)
            return fail(h, PTMC_E_ARG, "every run holds the potential's %d + %d atoms (got %d)", n.n1, n.n2, rh.d.n_atoms);
        rh.info.slot_base = slot;
        rh.info.local_base = 0;
        rh.kl = NNP2_NT;
        rh.lanes = NNP2_NT;
        rh.tile = -1;
        slot += rh.d.n_temps;
    }
    h->n_slots = slot;
    h->tiles.clear();
    h->launches.clear();
    Launch L;
    L.fn = (SweepFn)fs;
    L.init = (SweepFn)fi;
    L.ntiles = slot;
    L.wgs = NNP2_NT;
    L.kl = NNP2_NT;
    L.lds = h->nnp2_lds;
    h->launches.push_back(L);
    h->n_coord = (int64_t)slot * 3 * NNP_NP;
    h->n_comp = 0;
    int rc;
    if ((rc = upload(h, h->d_n2_rad, n.rad)) || (rc = upload(h, h->d_n2_ang, n.ang)) ||
        (rc = upload(h, h->d_n2_p1, n.p1)) || (rc = upload(h, h->d_n2_p2, n.p2)) ||
        (rc = upload(h, h->d_n2_grp, n.grp)) || (rc = upload(h, h->d_n2_gfn, n.gfn)))
        return rc;
    if ((rc = dalloc(h, h->d_n2_feat, (size_t)slot * NNP_NP * n.ng))) return rc;
    if ((rc = dalloc(h, h->d_n2_eatom, (size_t)slot * NNP_NP))) return rc;
    Nnp2Tables &t = h->nnp2_dev;
    t.n1 = n.n1;
    t.n_swaps = n.n_swaps;
    t.n_rad = n.n_rad;
    t.n_ang = n.n_ang;
    t.ng = n.ng;
    t.h1 = n.h1;
    t.h2 = n.h2;
    t.act1 = n.act[1];
    t.act2 = n.act[2];
    t.act3 = n.act[3];
    t.r_cut = n.r_cut;
    t.rad = h->d_n2_rad.p;
    t.ang = h->d_n2_ang.p;
    t.n_groups = n.n_groups;
    t.grp = h->d_n2_grp.p;
    t.gfn = h->d_n2_gfn.p;
    t.p1 = h->d_n2_p1.p;
    t.p2 = h->d_n2_p2.p;
    t.feat = h->d_n2_feat.p;
    t.eatom = h->d_n2_eatom.p;
    return PTMC_OK;
}

// Slots of a neural-network engine: one workgroup per slot, slots of a run
// consecutive, temperature t of run r starting in slot slot_base + t.
int prepare_nnp(ptmc_engine *h)
{
    if (h->cfg.precision != 1) return fail(h, PTMC_E_UNSUPPORTED, "the network potential runs in binary32 (precision 1)");
    if (h->cfg.ensemble != ENS_NVT || h->cfg.boundary != BC_SPHERE)
        return fail(h, PTMC_E_UNSUPPORTED, "the network potential runs NVT in a spherical container");
    int dev = 0, max_lds = 0;
    HIPCHK(h, hipGetDevice(&dev));
    HIPCHK(h, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // slots of 64 atoms, or 128 (one workgroup per CU) when a run is larger
    int n_max = 0;
    for (auto &rh : h->runs) n_max = std::max(n_max, rh.d.n_atoms);
    if (n_max > NNP_NP_WIDE)
        return fail(h, PTMC_E_UNSUPPORTED, "at most %d atoms with the network potential", NNP_NP_WIDE);
    h->nnp_np = n_max > NNP_NP ? NNP_NP_WIDE : NNP_NP;
    h->nnp_lds = nnp_lds_floats(h->nnp.nf, h->nnp.h1, h->nnp_np) * sizeof(float);
    if ((long long)h->nnp_lds > max_lds)
        return fail(h, PTMC_E_UNSUPPORTED, "network tables for %d-atom slots need %zu B of LDS (%d available)", h->nnp_np,
                    h->nnp_lds, max_lds);
    const void *fs = (const void *)nnp_sweep_fn(h->nnp_np), *fi = (const void *)nnp_init_fn(h->nnp_np);
    HIPCHK(h, hipFuncSetAttribute(fs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->nnp_lds));
    HIPCHK(h, hipFuncSetAttribute(fi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->nnp_lds));
    int slot = 0;
    for (auto &rh : h->runs) {
        rh.info.slot_base = slot;
        rh.info.local_base = 0;
        rh.kl = NNP_NT;
        rh.lanes = NNP_NT;
        rh.tile = -1;
        slot += rh.d.n_temps;
    }
    h->n_slots = slot;
    h->tiles.clear();
    h->launches.clear();
    Launch L;
    L.fn = (SweepFn)fs;
    L.init = (SweepFn)fi;
    L.ntiles = slot;
    L.wgs = NNP_NT;
    L.kl = NNP_NT;
    L.lds = h->nnp_lds;
    h->launches.push_back(L);
    h->n_coord = (int64_t)slot * 3 * h->nnp_np;
    h->n_comp = 0;
    int rc;
    auto &n = h->nnp;
    if ((rc = upload(h, h->d_nnp_rad, n.rad)) || (rc = upload(h, h->d_nnp_geta, n.geta)) ||
        (rc = upload(h, h->d_nnp_afn, n.afn)) || (rc = upload(h, h->d_nnp_rowoff, n.rowoff)) ||
        (rc = upload(h, h->d_nnp_w, n.w)) || (rc = upload(h, h->d_nnp_grp, n.grp)) ||
        (rc = upload(h, h->d_nnp_abuild, n.abuild)) || (rc = upload(h, h->d_nnp_gdesc, n.gdesc)) ||
        (rc = upload(h, h->d_nnp_wgroups, n.wgroups)))
        return rc;
    if ((rc = dalloc(h, h->d_feat, (size_t)slot * n.nf * h->nnp_np))) return rc;
    if ((rc = dalloc(h, h->d_eatom, (size_t)slot * h->nnp_np))) return rc;
    NnpTables &t = h->nnp_dev;
    t.n_rad = n.n_rad;
    t.n_ang = n.n_ang;
    t.nf = n.nf;
    t.h1 = n.h1;
    t.h2 = n.h2;
    t.n_rad_active = n.n_rad_active;
    t.n_groups = n.n_groups;
    t.r_cut = (float)n.r_cut;
    t.hardcore2 = n.hardcore2;
    t.rad = h->d_nnp_rad.p;
    t.grp = h->d_nnp_grp.p;
    t.grp_eta = h->d_nnp_geta.p;
    t.afn = h->d_nnp_afn.p;
    t.abuild = h->d_nnp_abuild.p;
    t.gdesc = h->d_nnp_gdesc.p;
    t.wgroups = h->d_nnp_wgroups.p;
    t.row_off = h->d_nnp_rowoff.p;
    t.w1 = h->d_nnp_w.p;
    t.b1 = t.w1 + (size_t)n.nf * n.h1;
    t.w2 = t.b1 + n.h1;
    t.b2 = t.w2 + (size_t)n.h1 * n.h2;
    t.w3 = t.b2 + n.h2;
    t.b3 = n.b3;
    return PTMC_OK;
}

template <typename Real>
void scatter_coords_nnp(ptmc_engine *h, std::vector<Real> &buf)
{
    buf.assign((size_t)h->n_coord, (Real)0);
    const int np = slot_np(h);
    for (auto &rh : h->runs) {
        const int N = rh.d.n_atoms;
        for (int t = 0; t < rh.d.n_temps; ++t) {
            const double *x = &rh.x0[(size_t)(t % rh.d.n_configs) * N * 3];
            Real *dst = &buf[(size_t)(rh.info.slot_base + t) * 3 * np];
            for (int j = 0; j < N; ++j)
                for (int c = 0; c < 3; ++c) dst[c * np + j] = (Real)x[3 * j + c];
        }
    }
}

int timing_event(ptmc_engine *h, hipEvent_t *ev)
{
    if (h->ev_free.empty()) {
        hipEvent_t e;
        HIPCHK(h, hipEventCreate(&e));
        h->ev_pool.push_back(e);
        *ev = e;
        return PTMC_OK;
    }
    *ev = h->ev_free.back();
    h->ev_free.pop_back();
    return PTMC_OK;
}

// one launch of group k on stream st; cycle launches are bracketed by
// timing events unless `timed` is false
int launch_one(ptmc_engine *h, size_t k, const SweepArgs &a, hipStream_t st, bool init, bool timed = true)
{
    const Launch &L = h->launches[k];
    SweepArgs b = a;
    b.tiles = h->d_tiles.p + L.tile0;
    void *args[] = {&b};
    if (init) {
        HIPCHK(h, hipLaunchKernel((const void *)L.init, dim3(L.ntiles), dim3(L.wgs), args, L.lds, st));
        return PTMC_OK;
    }
    if (!timed) {
        HIPCHK(h, hipLaunchKernel((const void *)L.fn, dim3(L.ntiles), dim3(L.wgs), args, L.lds, st));
        h->n_launch += 1;
        return PTMC_OK;
    }
    LaunchTiming lt{(int)k, nullptr, nullptr};
    int rc;
    if ((rc = timing_event(h, &lt.a)) || (rc = timing_event(h, &lt.b))) return rc;
    HIPCHK(h, hipEventRecord(lt.a, st));
    HIPCHK(h, hipLaunchKernel((const void *)L.fn, dim3(L.ntiles), dim3(L.wgs), args, L.lds, st));
    HIPCHK(h, hipEventRecord(lt.b, st));
    h->pending.push_back(lt);
    h->n_launch += 1;
    return PTMC_OK;
}

// after the engine stream has been synchronised: fold launch durations into
// the per-group sums and recycle the events
// fold the durations of the oldest `n` pending launches (all by default) into
// the per-group sums, waiting for each to finish, and recycle their events
int collect_timings(ptmc_engine *h, size_t n = (size_t)-1)
{
    n = std::min(n, h->pending.size());
    for (size_t k = 0; k < n; ++k) {
        const LaunchTiming &lt = h->pending[k];
        float ms = 0.f;
        HIPCHK(h, hipEventSynchronize(lt.b));
        HIPCHK(h, hipEventElapsedTime(&ms, lt.a, lt.b));
        h->launches[lt.group].ms += ms;
        h->launches[lt.group].n_done += 1;
        h->ev_free.push_back(lt.a);
        h->ev_free.push_back(lt.b);
    }
    h->pending.erase(h->pending.begin(), h->pending.begin() + (long)n);
    return PTMC_OK;
}

int launch_all(ptmc_engine *h, SweepArgs a, bool init = false)
{
    const size_t nl = h->launches.size();
    if (nl == 1) return launch_one(h, 0, a, h->stream, init);
    // launch groups are independent: fork them onto side streams, join back
    HIPCHK(h, hipEventRecord(h->fork, h->stream));
    for (size_t k = 0; k < nl; ++k) {
        hipStream_t st = h->side[k % h->side.size()];
        HIPCHK(h, hipStreamWaitEvent(st, h->fork, 0));
        int rc = launch_one(h, k, a, st, init);
        if (rc) return rc;
        HIPCHK(h, hipEventRecord(h->join[k], st));
    }
    for (size_t k = 0; k < nl; ++k) HIPCHK(h, hipStreamWaitEvent(h->stream, h->join[k], 0));
    return PTMC_OK;
}

}  // namespace

extern "C" {

int ptmc_create(const ptmc_config *cfg, ptmc_handle *out)
{
    if (!cfg || !out) return PTMC_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PTMC_E_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return PTMC_E_ARG;
    if (cfg->precision < 0 || cfg->precision > 1 || cfg->ensemble < 0 || cfg->ensemble > 3 ||
        cfg->boundary < 0 || cfg->boundary > 3 || cfg->exchange_mode < 0 || cfg->exchange_mode > 2 ||
        cfg->n_adjust <= 0 || cfg->mc_sample <= 0 || cfg->n_bin < 2)
        return PTMC_E_ARG;
    if ((cfg->ensemble == 1 || cfg->ensemble == 2) && cfg->boundary == BC_SPHERE) return PTMC_E_ARG;  // src/Ensembles.jl:191-193
    // separated volume moves scale x, y and z apart: boxes with a height only
    // (scale_xy / scale_z have no CubicBC method, src/MCMoves.jl:189-225)
    if (cfg->ensemble == 2 && cfg->boundary != BC_RECT && cfg->boundary != BC_RHOMB) return PTMC_E_ARG;
    ptmc_engine *h = new ptmc_engine();
    h->cfg = *cfg;
    h->separated = cfg->ensemble == 2;
    if (h->separated) h->cfg.ensemble = ENS_NPT;
    // NNVT: atom moves plus atom swaps at fixed volume (src/Ensembles.jl:138-150)
    h->nnvt = cfg->ensemble == 3;
    if (h->nnvt) h->cfg.ensemble = ENS_NVT;
    if (h->cfg.cycles_per_launch <= 0) h->cfg.cycles_per_launch = 256;
    if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        delete h;
        return PTMC_E_DEVICE;
    }
    *out = h;
    return PTMC_OK;
}

int ptmc_destroy(ptmc_handle h)
{
    if (!h) return PTMC_E_ARG;
    (void)hipSetDevice(h->cfg.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->d_tiles.free_();
    h->d_runs.free_();
    h->d_run_list.free_();
    h->d_slot_run.free_();
    h->d_slot_temp.free_();
    h->d_run_error.free_();
    h->d_slot_en.free_();
    h->d_slot_box.free_();
    h->d_slot_boxh.free_();
    h->d_lut64.free_();
    h->d_lut32.free_();
    h->d_plug_rad.free_();
    h->d_plug_ang.free_();
    h->d_plug_w.free_();
    h->d_plug.free_();
    h->d_maxboxh.free_();
    h->d_lh.free_();
    h->d_displ_sep.free_();
    h->d_count_volxy.free_();
    h->d_count_volz.free_();
    h->d_slot_drift.free_();
    h->d_coords.free_();
    h->d_comps.free_();
    h->d_nnp_rad.free_();
    h->d_nnp_geta.free_();
    h->d_nnp_afn.free_();
    h->d_nnp_rowoff.free_();
    h->d_nnp_w.free_();
    h->d_feat.free_();
    h->d_eatom.free_();
    h->d_nnp_grp.free_();
    h->d_nnp_abuild.free_();
    h->d_nnp_gdesc.free_();
    h->d_nnp_wgroups.free_();
    h->d_temp_slot.free_();
    h->d_n2_rad.free_();
    h->d_n2_ang.free_();
    h->d_n2_p1.free_();
    h->d_n2_p2.free_();
    h->d_n2_feat.free_();
    h->d_n2_eatom.free_();
    h->d_n2_grp.free_();
    h->d_n2_gfn.free_();
    h->d_sur_w.free_();
    h->d_sur_m.free_();
    h->d_sur_v.free_();
    h->d_sur_grad.free_();
    h->d_sur_partial.free_();
    h->d_sur_D.free_();
    h->d_sur_x.free_();
    h->d_sur_E.free_();
    h->d_sur_pred.free_();
    h->d_sur_N.free_();
    h->d_sur_idx.free_();
    h->d_da_stats.free_();
    h->d_beta.free_();
    h->d_displ.free_();
    h->d_maxbox.free_();
    h->d_ham.free_();
    h->d_count_atom.free_();
    h->d_count_vol.free_();
    h->d_count_exc.free_();
    h->d_hist.free_();
    h->d_trace_de.free_();
    h->d_trace_acc.free_();
    h->d_trace_idx.free_();
    h->d_trace_aux.free_();
    for (auto st : h->side) (void)hipStreamSynchronize(st);
    for (auto ev : h->join) (void)hipEventDestroy(ev);
    for (auto ev : h->ev_pool) (void)hipEventDestroy(ev);
    if (h->fork) (void)hipEventDestroy(h->fork);
    for (auto st : h->side) (void)hipStreamDestroy(st);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PTMC_OK;
}

int ptmc_last_error(ptmc_handle h, char *buf, size_t len)
{
    if (!h || !buf || len == 0) return PTMC_E_ARG;
    snprintf(buf, len, "%s", h->err.c_str());
    return PTMC_OK;
}

int ptmc_set_potential_elj(ptmc_handle h, int32_t even_powers, int32_t n_coeff, const double *coeff)
{
    if (!h || !coeff || n_coeff <= 0 || n_coeff > MAX_COEFF) return h ? fail(h, PTMC_E_ARG, "bad ELJ coefficients") : PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "potential must be set before ptmc_prepare");
    PotParams p{};
    p.kind = even_powers ? POT_ELJ_EVEN : POT_ELJ;
    p.ncoeff = n_coeff;
    for (int i = 0; i < n_coeff; ++i) p.coeff[i] = coeff[i];
    h->pot = p;
    h->pot_set = true;
    return PTMC_OK;
}

int ptmc_set_potential_eljb(ptmc_handle h, int32_t n_coeff, const double *coeff_c, const double *coeff_a,
                            const double *coeff_b)
{
    if (!h || !coeff_c || !coeff_a || !coeff_b || n_coeff <= 0 || n_coeff > 8)
        return h ? fail(h, PTMC_E_ARG, "bad ELJPotentialB coefficients (1 to 8 of each)") : PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "potential must be set before ptmc_prepare");
    PotParams p{};
    p.kind = POT_ELJB;
    p.ncoeff = n_coeff;
    for (int i = 0; i < n_coeff; ++i) {
        p.coeff[i] = coeff_c[i];
        p.coeff_a[i] = coeff_a[i];
        p.coeff_b[i] = coeff_b[i];
    }
    h->pot = p;
    h->pot_set = true;
    return PTMC_OK;
}

int ptmc_set_potential_lookup(ptmc_handle h, int32_t n_dist, int32_t n_angle, double start_dist, double d_dist,
                              double d_angle, double c6coeff, const double *table)
{
    if (!h) return PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "potential must be set before ptmc_prepare");
    if (!table || n_dist < 1 || n_angle < 91 || !(d_dist > 0) || !(d_angle > 0) || !(start_dist >= 0) ||
        !std::isfinite(c6coeff))
        return fail(h, PTMC_E_ARG, "bad lookup table (n_angle >= 91 rows, positive steps)");
    // the largest angle index the lookup forms, at 89.5 degrees, must be a row
    if (std::rint(89.5 / d_angle + 1.0) > (double)n_angle)
        return fail(h, PTMC_E_ARG, "angle step %g needs %d rows", d_angle, (int)std::rint(89.5 / d_angle + 1.0));
    const size_t n = (size_t)n_dist * n_angle;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(table[i])) return fail(h, PTMC_E_ARG, "lookup table entry %zu is not finite", i);
    PotParams p{};
    p.kind = POT_LUT;
    p.ncoeff = 1;
    p.coeff[0] = c6coeff;  // the tail correction reads c6 from the coefficient slot
    p.lut_ndist = n_dist;
    p.lut_nangle = n_angle;
    p.lut_start = start_dist;
    p.lut_d = d_dist;
    p.lut_dangle = d_angle;
    p.lut_c6 = c6coeff;
    h->lut_host.assign(table, table + n);
    h->pot = p;
    h->pot_set = true;
    return PTMC_OK;
}

int ptmc_set_potential_eam(ptmc_handle h, double n, double m, double ean, double ecam)
{
    if (!h) return PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "potential must be set before ptmc_prepare");
    if (!(n > 0) || !(m > 0) || !std::isfinite(ean) || !std::isfinite(ecam))
        return fail(h, PTMC_E_ARG, "bad EAM parameters");
    PotParams p{};
    p.kind = POT_EAM;
    p.eam_n = n;
    p.eam_m = m;
    p.eam_ean = ean;
    p.eam_ecam = ecam;
    h->pot = p;
    h->pot_set = true;
    return PTMC_OK;
}

int ptmc_set_potential_nnp2(ptmc_handle h, int32_t n1, int32_t n2, int32_t n_swaps, int32_t n_rad, int32_t n_ang,
                            const double *rad, const double *ang, double r_cut, int32_t n_layers, const int32_t *nodes,
                            const int32_t *activations, int32_t n_params, const double *params1,
                            const double *params2)
{
    if (!h) return PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "potential must be set before ptmc_prepare");
    if (!h->nnvt) return fail(h, PTMC_E_UNSUPPORTED, "the two-species network potential runs in the NNVT ensemble (3)");
    if (n1 < 0 || n2 < 0 || n1 + n2 < 2 || n1 + n2 > NNP_NP || n_swaps < 0 || (n_swaps > 0 && (n1 < 1 || n2 < 1)))
        return fail(h, PTMC_E_ARG, "bad species counts (%d, %d, at most %d atoms) or swaps per cycle (%d)", n1, n2,
                    NNP_NP, n_swaps);
    if (n_rad < 0 || n_ang < 0 || n_rad + n_ang < 1 || (n_rad && !rad) || (n_ang && !ang) || !(r_cut > 0) ||
        !nodes || !activations || !params1 || !params2)
        return fail(h, PTMC_E_ARG, "bad network potential description");
    const int ng = 2 * n_rad + 3 * n_ang;
    if (n_layers != 4 || nodes[0] != ng || nodes[3] != 1 || nodes[1] < 1 || nodes[2] < 1)
        return fail(h, PTMC_E_UNSUPPORTED, "network shape must be [2 n_rad + 3 n_ang, h1, h2, 1]");
    for (int l = 1; l < 4; ++l)
        if (activations[l] != 1 && activations[l] != 2)
            return fail(h, PTMC_E_UNSUPPORTED, "activations are 1 (linear) or 2 (tanh)");
    const int h1 = nodes[1], h2 = nodes[2];
    if (n_params != ng * h1 + h1 + h1 * h2 + h2 + h2 + 1)
        return fail(h, PTMC_E_ARG, "parameter count does not match the layers");
    for (int q = 0; q < n_ang; ++q) {
        const double zeta = ang[(size_t)q * 15 + 2], base_lo = 1.0 - fabs(ang[(size_t)q * 15 + 1]);
        if (!(zeta >= 0.0) || (base_lo < 0.0 && zeta != rint(zeta)))
            return fail(h, PTMC_E_ARG, "angular function %d: zeta must be non-negative (integer when |lambda| > 1)", q);
    }
    auto &n = h->nnp2;
    n.n1 = n1;
    n.n2 = n2;
    n.n_swaps = n_swaps;
    n.n_rad = n_rad;
    n.n_ang = n_ang;
    n.ng = ng;
    n.h1 = h1;
    n.h2 = h2;
    for (int l = 0; l < 4; ++l) n.act[l] = activations[l];
    n.r_cut = r_cut;
    n.rad.assign(rad, rad + (size_t)n_rad * 9);
    n.ang.assign(ang, ang + (size_t)n_ang * 15);
    if (n.rad.empty()) n.rad.assign(9, 0.0);
    if (n.ang.empty()) n.ang.assign(15, 0.0);
    n.p1.assign(params1, params1 + n_params);
    n.p2.assign(params2, params2 + n_params);
    // angular functions grouped by equal eta (first appearance order, stable
    // inside a group), at most NNP2_GMAX per group
    n.grp.clear();
    n.gfn.clear();
    {
        std::vector<char> done(n_ang, 0);
        for (int q = 0; q < n_ang; ++q) {
            if (done[q]) continue;
            const double eta = ang[(size_t)q * 15];
            int cnt = 0;
            for (int r = q; r < n_ang; ++r) {
                if (done[r] || ang[(size_t)r * 15] != eta) continue;
                if (cnt == NNP2_GMAX) {
                    n.grp.push_back((int32_t)n.gfn.size() - cnt);
                    n.grp.push_back(cnt);
                    cnt = 0;
                }
                n.gfn.push_back(r);
                done[r] = 1;
                ++cnt;
            }
            n.grp.push_back((int32_t)n.gfn.size() - cnt);
            n.grp.push_back(cnt);
        }
        n.n_groups = (int)n.grp.size() / 2;
        if (n.gfn.empty()) n.gfn.assign(1, 0);
        if (n.grp.empty()) n.grp.assign(2, 0);
    }
    PotParams p{};
    p.kind = POT_NNP2;
    h->pot = p;
    h->pot_set = true;
    return PTMC_OK;
}

int ptmc_set_potential_nnp(ptmc_handle h, int32_t n_rad, int32_t n_ang, const double *rad, const double *ang,
                           double r_cut, double hardcore2, int32_t q10_compat, int32_t n_layers, const int32_t *nodes,
                           const int32_t *activations, int32_t n_params, const double *params)
{
    if (!h) return PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "potential must be set before ptmc_prepare");
    if (n_rad < 0 || n_ang < 0 || n_rad + n_ang < 1 || (n_rad && !rad) || (n_ang && !ang) || !(r_cut > 0) ||
        !nodes || !activations || !params)
        return fail(h, PTMC_E_ARG, "bad network potential description");
    if (n_layers != 4 || nodes[0] != n_rad + n_ang || nodes[3] != 1 || nodes[1] < 1 || nodes[1] > NNP_MAXH ||
        nodes[2] < 1 || nodes[2] > NNP_MAXH || activations[1] != 2 || activations[2] != 2 || activations[3] != 1)
        return fail(h, PTMC_E_UNSUPPORTED, "network shape must be [n_features, h1 <= %d (tanh), h2 <= %d (tanh), 1 (linear)]",
                    NNP_MAXH, NNP_MAXH);
    const int nf = n_rad + n_ang, h1 = nodes[1], h2 = nodes[2];
    if (nf > NNP_MAXF) return fail(h, PTMC_E_UNSUPPORTED, "at most %d symmetry functions", NNP_MAXF);
    if (n_params != nf * h1 + h1 + h1 * h2 + h2 + h2 + 1) return fail(h, PTMC_E_ARG, "parameter count does not match the layers");
    auto &n = h->nnp;
    n = {};
    n.n_rad = n_rad;
    n.n_ang = n_ang;
    n.nf = nf;
    n.h1 = h1;
    n.h2 = h2;
    n.r_cut = r_cut;
    n.hardcore2 = hardcore2;
    n.q10 = q10_compat ? 1 : 0;
    n.rowoff.assign(nf, 0.0f);
    for (int q = 0; q < n_rad; ++q) {
        const double *r = rad + 4 * q;
        if (r[3] == 0.0) continue;
        n.rad.insert(n.rad.end(), {(float)r[0], (float)r[1], (float)q, 0.0f});
        n.rowoff[q] = (float)r[2];
        n.n_rad_active += 1;
    }
    // Angular functions grouped by eta (stable order), at most NNP_GMAX per
    // group; inside a group lambda = +1 comes before -1 and zeta ascends, so
    // the powers (1 + lambda cos)^zeta of a group follow from one chain of
    // squarings per lambda.  On the device the feature rows of a group are
    // consecutive (row = n_rad + table index; inactive functions follow), and
    // the first-layer weights and the row offsets are permuted to match.
    const int n_built = q10_compat ? std::max(0, n_ang - n_rad) : n_ang;
    auto zeta_code = [&](int q) {
        for (int e = 0; e < NNP_KSLOT; ++e)
            if (ang[6 * q + 2] == (double)(1 << e)) return e;
        return -1;
    };
    std::vector<int> idx;
    for (int q = 0; q < n_ang; ++q) {
        if (ang[6 * q + 5] == 0.0) continue;
        if (zeta_code(q) < 0)
            return fail(h, PTMC_E_UNSUPPORTED, "angular function %d: zeta = %g is not a power of two up to %d", q,
                        ang[6 * q + 2], 1 << (NNP_KSLOT - 1));
        if (ang[6 * q + 1] != 1.0 && ang[6 * q + 1] != -1.0)
            return fail(h, PTMC_E_UNSUPPORTED, "angular function %d: lambda = %g is not +-1", q, ang[6 * q + 1]);
        idx.push_back(q);
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return ang[6 * x] < ang[6 * y]; });
    std::vector<int> dev_row(nf, -1);  // original feature row -> device row
    for (int q = 0; q < n_rad; ++q) dev_row[q] = q;
    auto slot_of = [&](int q) { return (ang[6 * q + 1] > 0.0 ? 0 : NNP_KSLOT) + zeta_code(q); };
    for (size_t k = 0; k < idx.size();) {
        const double eta = ang[6 * idx[k]];
        const size_t k0 = k;
        uint32_t used = 0;  // one function per (lambda, k) slot in a group
        while (k < idx.size() && ang[6 * idx[k]] == eta && !(used >> slot_of(idx[k]) & 1u)) used |= 1u << slot_of(idx[k++]);
        std::stable_sort(idx.begin() + k0, idx.begin() + k, [&](int x, int y) { return slot_of(x) < slot_of(y); });
        const int first = (int)(n.afn.size() / 4);
        std::vector<int32_t> d(NNP_GD, 0);
        const float eta2 = (float)(eta * 1.4426950408889634);  // e^{-eta x} = 2^{-eta log2(e) x}
        std::memcpy(&d[0], &eta2, 4);
        d[1] = n_rad + first;
        d[2] = d[3] = -1;
        for (int sl = 0; sl < NNP_GMAX; ++sl) d[4 + NNP_GMAX + sl] = -1;
        for (size_t e = k0; e < k; ++e) {
            const int q = idx[e];
            const double *a = ang + 6 * q;
            const int row = n_rad + (int)(n.afn.size() / 4);
            dev_row[n_rad + q] = row;
            // zeta = 2^k is stored as k (k squarings)
            n.afn.insert(n.afn.end(), {(float)a[1], (float)zeta_code(q), (float)a[3], (float)row});
            n.abuild.push_back(q < n_built ? 1 : 0);
            if (q < n_built) n.rowoff[row] = (float)a[4];
            const int sl = slot_of(q);
            const float tpz = (float)a[3];
            std::memcpy(&d[4 + sl], &tpz, 4);
            d[4 + NNP_GMAX + sl] = row - d[1];
            int &kmax = d[sl < NNP_KSLOT ? 2 : 3];
            kmax = std::max(kmax, zeta_code(q));
        }
        n.gdesc.insert(n.gdesc.end(), d.begin(), d.end());
        n.grp.insert(n.grp.end(), {first, (int)(k - k0)});
        n.geta.push_back((float)eta);
        n.n_groups += 1;
    }
    if (n.n_groups > NNP_MAXG)
        return fail(h, PTMC_E_UNSUPPORTED, "%d angular groups (distinct eta, or repeated (lambda, zeta) within one eta); "
                    "at most %d", n.n_groups, NNP_MAXG);
    // groups dealt to the waves of the move update (each wave owns its groups'
    // rows): largest work first onto the least loaded wave, at most NNP_WG each
    {
        constexpr int NW = NNP_NT / 64;
        std::vector<int> order(n.n_groups);
        std::vector<double> cost(n.n_groups), load(NW, 0.0);
        std::vector<std::vector<int>> lists(NW);
        for (int g = 0; g < n.n_groups; ++g) {
            const int32_t *d = &n.gdesc[(size_t)g * NNP_GD];
            order[g] = g;
            cost[g] = 2.0 + 2.0 * ((d[2] + 1) + (d[3] + 1));  // exps + per-slot chain work
        }
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
        for (int g : order) {
            int best = -1;
            for (int w = 0; w < NW; ++w)
                if ((int)lists[w].size() < NNP_WG && (best < 0 || load[w] < load[best])) best = w;
            lists[best].push_back(g);
            load[best] += cost[g];
        }
        n.wgroups.assign((size_t)NW * (NNP_WG + 1), 0);
        for (int w = 0; w < NW; ++w) {
            n.wgroups[(size_t)w * (NNP_WG + 1)] = (int32_t)lists[w].size();
            for (size_t e = 0; e < lists[w].size(); ++e) n.wgroups[(size_t)w * (NNP_WG + 1) + 1 + e] = lists[w][e];
        }
    }
    int next = n_rad + (int)(n.afn.size() / 4);
    for (int q = 0; q < n_ang; ++q)
        if (dev_row[n_rad + q] < 0) dev_row[n_rad + q] = next++;
    n.w.resize((size_t)n_params - 1);
    for (int k = 0; k + 1 < n_params; ++k) n.w[k] = (float)params[k];
    for (int q = 0; q < nf; ++q)
        for (int j = 0; j < h1; ++j) n.w[(size_t)dev_row[q] * h1 + j] = (float)params[(size_t)q * h1 + j];
    n.b3 = (float)params[n_params - 1];
    n.rad64.assign(rad, rad + (size_t)4 * n_rad);
    n.ang64.assign(ang, ang + (size_t)6 * n_ang);
    n.w64.assign(params, params + n_params);
    h->plug_nnp_up = false;
    PotParams p{};
    p.kind = POT_NNP;
    h->pot = p;
    h->pot_set = true;
    return PTMC_OK;
}

int ptmc_add_run(ptmc_handle h, const ptmc_run_desc *d, int32_t *run_index)
{
    if (!h || !d) return PTMC_E_ARG;
    if (h->prepared) return fail(h, PTMC_E_STATE, "runs must be added before ptmc_prepare");
    if (d->n_atoms < 2 || d->n_temps < 1 || !d->beta || !d->x0 || d->n_configs < 1)
        return fail(h, PTMC_E_ARG, "bad run description");
    if (h->cfg.boundary == BC_SPHERE && !(d->radius > 0)) return fail(h, PTMC_E_ARG, "radius must be positive");
    if (h->cfg.boundary != BC_SPHERE && !(d->box_length > 0)) return fail(h, PTMC_E_ARG, "box length must be positive");
    if ((h->cfg.boundary == BC_RECT || h->cfg.boundary == BC_RHOMB) && !(d->box_height > 0))
        return fail(h, PTMC_E_ARG, "box height must be positive");
    if (d->n_temps > (1 << 20)) return fail(h, PTMC_E_ARG, "too many temperatures");
    // 32 and 64 lanes per configuration exist in the spread layout only; the
    // ladder layout's planner refuses them as unsupported
    if (d->lanes_per_replica < 0 || d->lanes_per_replica > 64 || (d->lanes_per_replica & (d->lanes_per_replica - 1)))
        return fail(h, PTMC_E_ARG, "lanes_per_replica must be 0 or a power of two up to 64");
    RunHost rh;
    rh.d = *d;
    rh.beta.assign(d->beta, d->beta + d->n_temps);
    rh.x0.assign(d->x0, d->x0 + (size_t)d->n_configs * d->n_atoms * 3);
    rh.d.beta = nullptr;
    rh.d.x0 = nullptr;
    if (run_index) *run_index = (int32_t)h->runs.size();
    h->runs.push_back(std::move(rh));
    return PTMC_OK;
}

int ptmc_prepare(ptmc_handle h)
{
    if (!h) return PTMC_E_ARG;
    if (h->prepared) return PTMC_OK;
    if (!h->pot_set) return fail(h, PTMC_E_STATE, "no potential set");
    if (h->runs.empty()) return fail(h, PTMC_E_STATE, "no runs added");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    if (h->nnvt != is_nnp2(h))
        return fail(h, PTMC_E_UNSUPPORTED, "the NNVT ensemble and the two-species network potential go together");
    int rc = is_nnp(h) ? prepare_nnp(h) : (is_nnp2(h) ? prepare_nnp2(h) : choose_pair(h));
    if (rc) return rc;
    if (is_pair_slot(h) && (rc = prepare_pair_slot(h))) return rc;
    if (!is_slot(h)) rc = plan(h);
    if (rc) return rc;
    if (h->launches.size() > 1) {
        const size_t ns = std::min<size_t>(h->launches.size(), 4);
        while (h->side.size() < ns) {
            hipStream_t st;
            HIPCHK(h, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            h->side.push_back(st);
        }
        if (!h->fork) HIPCHK(h, hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
        while (h->join.size() < h->launches.size()) {
            hipEvent_t ev;
            HIPCHK(h, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            h->join.push_back(ev);
        }
    }

    const int nb2 = h->cfg.n_bin + 2;
    h->hist_stride = h->cfg.ensemble == ENS_NPT ? nb2 * nb2 : nb2;

    std::vector<RunInfo> infos;
    std::vector<int32_t> slot_run(h->n_slots), slot_temp(h->n_slots);
    std::vector<double> slot_en(h->n_slots, 0.0), slot_box(h->n_slots, 1.0), slot_boxh(h->n_slots, 1.0);
    std::vector<double> beta, displ, maxbox, maxboxh, lh, displ_sep;
    const bool pbc = h->cfg.boundary != BC_SPHERE;
    const bool has_h = h->cfg.boundary == BC_RECT || h->cfg.boundary == BC_RHOMB;
    int tb = 0;
    for (int r = 0; r < (int)h->runs.size(); ++r) {
        RunHost &rh = h->runs[r];
        RunInfo &ri = rh.info;
        ri.n_atoms = rh.d.n_atoms;
        ri.n_temps = rh.d.n_temps;
        ri.temp_base = tb;
        ri.run_key = rh.d.run_key == 0xffffffffu ? (uint32_t)r : rh.d.run_key;
        ri.radius2 = rh.d.radius * rh.d.radius;  // SphericalBC stores radius^2, src/BoundaryConditions.jl:44-47
        ri.pressure = rh.d.pressure;
        ri.en_min = 0.0;
        ri.delta_en = 1.0;
        ri.v_min = 0.0;
        ri.delta_v = 1.0;
        infos.push_back(ri);
        for (int t = 0; t < rh.d.n_temps; ++t) {
            slot_run[ri.slot_base + t] = r;
            slot_temp[ri.slot_base + t] = t;
            const double bl = pbc ? rh.d.box_length : 1.0, bh = has_h ? rh.d.box_height : bl;
            slot_box[ri.slot_base + t] = bl;
            slot_boxh[ri.slot_base + t] = bh;
            beta.push_back(rh.beta[t]);
            displ.push_back(rh.d.max_displ_atom);
            displ.push_back(rh.d.max_displ_vol);
            // max_length = 1.8 L, max_height = 1.8 H, lh_ratio = max_length / max_height
            // (src/MCStates.jl:63-86, :106; the cubic box's height is its length)
            maxbox.push_back(pbc ? bl * 1.8 : 30.0);
            maxboxh.push_back(pbc ? bh * 1.8 : 30.0);
            lh.push_back(maxbox.back() / maxboxh.back());
            displ_sep.push_back(rh.d.max_displ_vol);  // max_displ[3] (xy) and [4] (z): the reference default 0.1 each
            displ_sep.push_back(rh.d.max_displ_vol);
        }
        tb += rh.d.n_temps;
    }
    h->n_temps = tb;
    {
        std::vector<int32_t> ts(tb);
        for (auto &rh : h->runs)
            for (int t = 0; t < rh.d.n_temps; ++t) ts[rh.info.temp_base + t] = rh.info.slot_base + t;
        if ((rc = upload(h, h->d_temp_slot, ts))) return rc;
    }

    if ((rc = upload(h, h->d_tiles, h->tiles))) return rc;
    if ((rc = upload(h, h->d_runs, infos))) return rc;
    if ((rc = upload(h, h->d_slot_run, slot_run))) return rc;
    if ((rc = upload(h, h->d_slot_temp, slot_temp))) return rc;
    if ((rc = upload(h, h->d_slot_en, slot_en))) return rc;
    if ((rc = upload(h, h->d_slot_box, slot_box))) return rc;
    if ((rc = upload(h, h->d_slot_boxh, slot_boxh))) return rc;
    if ((rc = dalloc(h, h->d_slot_drift, h->n_slots))) return rc;
    if ((rc = dalloc(h, h->d_run_error, h->runs.size()))) return rc;
    if ((rc = upload(h, h->d_beta, beta))) return rc;
    if ((rc = upload(h, h->d_displ, displ))) return rc;
    if ((rc = upload(h, h->d_maxbox, maxbox))) return rc;
    if ((rc = upload(h, h->d_maxboxh, maxboxh))) return rc;
    if ((rc = upload(h, h->d_lh, lh))) return rc;
    if ((rc = upload(h, h->d_displ_sep, displ_sep))) return rc;
    if ((rc = dalloc(h, h->d_count_volxy, (size_t)2 * tb))) return rc;
    if ((rc = dalloc(h, h->d_count_volz, (size_t)2 * tb))) return rc;
    if ((rc = dalloc(h, h->d_ham, (size_t)2 * tb))) return rc;
    if ((rc = dalloc(h, h->d_count_atom, (size_t)2 * tb))) return rc;
    if ((rc = dalloc(h, h->d_count_vol, (size_t)2 * tb))) return rc;
    if ((rc = dalloc(h, h->d_count_exc, (size_t)2 * tb))) return rc;
    if ((rc = dalloc(h, h->d_hist, (size_t)h->hist_stride * tb))) return rc;
    if (slot_f32(h)) {
        std::vector<float> buf;
        scatter_coords_nnp(h, buf);
        std::vector<unsigned char> raw(buf.size() * sizeof(float));
        memcpy(raw.data(), buf.data(), raw.size());
        if ((rc = upload(h, h->d_coords, raw))) return rc;
    } else if (is_nnp2(h)) {
        std::vector<double> buf;
        scatter_coords_nnp(h, buf);
        std::vector<unsigned char> raw(buf.size() * sizeof(double));
        memcpy(raw.data(), buf.data(), raw.size());
        if ((rc = upload(h, h->d_coords, raw))) return rc;
    } else if (h->cfg.precision == 0) {
        std::vector<double> buf;
        scatter_coords<double>(h, buf);
        std::vector<unsigned char> raw(buf.size() * sizeof(double));
        memcpy(raw.data(), buf.data(), raw.size());
        if ((rc = upload(h, h->d_coords, raw))) return rc;
    } else {
        std::vector<float> buf;
        scatter_coords<float>(h, buf);
        std::vector<unsigned char> raw(buf.size() * sizeof(float));
        memcpy(raw.data(), buf.data(), raw.size());
        if ((rc = upload(h, h->d_coords, raw))) return rc;
    }
    if ((rc = dalloc(h, h->d_comps, (size_t)h->n_comp * (h->cfg.precision == 0 ? 8 : 4)))) return rc;
    h->prepared = true;

    // initial energies (initialise_energy) on the device in FP64
    SweepArgs a = base_args(h);
    a.init_run = -1;
    if ((rc = launch_all(h, a, true))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return PTMC_OK;
}

int ptmc_start_phase(ptmc_handle h, int32_t production)
{
    if (!h) return PTMC_E_ARG;
    h->phase_counter = 0;
    h->production = production ? 1 : 0;
    return PTMC_OK;
}

int ptmc_cycles(ptmc_handle h, int64_t n_cycles)
{
    if (!h) return PTMC_E_ARG;
    if (n_cycles < 0) return fail(h, PTMC_E_ARG, "negative cycle count");
    int rc = ptmc_prepare(h);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int64_t done = 0;
    // spread layout: whether the side streams must wait for work pending on
    // the engine stream (set-up copies, counter resets), asked before the
    // start event goes onto it
    const bool engine_busy = h->spread && h->launches.size() > 1 && hipStreamQuery(h->stream) != hipSuccess;
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    if (is_slot(h)) {
        // one launch per cycle, then exchange / adaptation / sampling
        const bool tracing = h->trace_pending > 0;
        if (tracing && n_cycles > h->trace_cycles)
            return fail(h, PTMC_E_STATE, "traced cycles must be run in one ptmc_cycles call of at most %d cycles",
                        h->trace_cycles);
        h->trace_pending = 0;
        int n_runs = (int)h->runs.size();
        for (int64_t c = 0; c < n_cycles; ++c) {
            SweepArgs a = base_args(h);
            a.sweep0 = h->sweep_counter;
            a.cycle0 = h->phase_counter;
            a.n_cycles = 1;
            a.sampling = h->production;
            a.init_run = -1;
            if (tracing) {
                a.trace_de = h->d_trace_de.p;
                a.trace_acc = h->d_trace_acc.p;
                a.trace_idx = h->d_trace_idx.p;
                a.trace_aux = h->d_trace_aux.p;
                a.trace_cycles = h->trace_cycles;
                a.trace_cycle = (int32_t)c;
            }
            if ((rc = launch_one(h, 0, a, h->stream, false))) return rc;
            void *args[] = {&a, &n_runs};
            HIPCHK(h, hipLaunchKernel((const void *)nnp_cycle_fn(), dim3((n_runs + 63) / 64), dim3(64), args, 0,
                                      h->stream));
            if (h->sur.init && h->sur.collect_every > 0 && (h->sweep_counter + 1) % h->sur.collect_every == 0) {
                HIPCHK(h, sur_launch_collect(a, h->n_slots, h->d_sur_D.p, h->d_sur_E.p, h->d_sur_N.p, h->sur.n_samples,
                                             h->sur.cap, h->stream));
                h->sur.n_samples += h->n_slots;
            }
            if (h->pending.size() >= 512 && (rc = collect_timings(h, 256))) return rc;
            h->sweep_counter += 1;
            h->phase_counter += 1;
        }
        done = n_cycles;
    }
    if (h->spread) {
        // one launch of the tiles runs the cycles up to the next exchange
        // (exchange_every; one cycle with the default), then the per-ladder
        // exchange, adaptation and sampling of that cycle; a last launch
        // without moves applies a recompute due after the final cycle
        const bool tracing = h->trace_pending > 0;
        if (tracing && n_cycles > h->trace_cycles)
            return fail(h, PTMC_E_STATE, "traced cycles must be run in one ptmc_cycles call of at most %d cycles",
                        h->trace_cycles);
        h->trace_pending = 0;
        const size_t ng = h->launches.size();
        // groups on their own streams (one group: the engine stream), joined at the end
        auto stream_of = [&](size_t g) { return ng == 1 ? h->stream : h->side[g % h->side.size()]; };
        // the side streams start after the engine stream's pending work;
        // nothing to wait for when it was idle
        if (ng > 1 && engine_busy) {
            HIPCHK(h, hipEventRecord(h->fork, h->stream));
            for (size_t g = 0; g < ng; ++g) HIPCHK(h, hipStreamWaitEvent(stream_of(g), h->fork, 0));
        }
        const int64_t k = h->cfg.exchange_every > 1 ? h->cfg.exchange_every : 1;
        for (int64_t c = 0; c < n_cycles;) {
            // cycles up to and including the next with (sweep + 1) % k == 0
            const int64_t n = std::min<int64_t>(k - h->sweep_counter % k, n_cycles - c);
            const bool exch = (h->sweep_counter + n) % k == 0;
            auto args_of = [&](size_t g) {
                const Launch &L = h->launches[g];
                SweepArgs a = base_args(h);
                a.sweep0 = h->sweep_counter;
                a.cycle0 = h->phase_counter;
                a.n_cycles = (int32_t)n;
                a.sampling = h->production;
                a.init_run = -1;
                a.run_list = h->d_run_list.p + L.run0;
                a.recompute_done = h->recompute_applied == h->sweep_counter ? 1 : 0;
                if (tracing) {
                    a.trace_de = h->d_trace_de.p;
                    a.trace_acc = h->d_trace_acc.p;
                    a.trace_idx = h->d_trace_idx.p;
                    a.trace_cycles = h->trace_cycles;
                    a.trace_cycle = (int32_t)c;
                }
                return a;
            };
            // every group's moves first, so the groups start together
            const bool timed = h->spread_launch++ % SPREAD_TIME_EVERY == 0;
            for (size_t g = 0; g < ng; ++g)
                if ((rc = launch_one(h, g, args_of(g), stream_of(g), false, timed))) return rc;
            if (exch)
                for (size_t g = 0; g < ng; ++g) {
                    // the exchange cycle is the launch's last
                    SweepArgs a = args_of(g);
                    a.sweep0 = h->sweep_counter + n - 1;
                    a.cycle0 = h->phase_counter + n - 1;
                    void *args[] = {&a};
                    HIPCHK(h, hipLaunchKernel(spread_cycle_fn(), dim3(h->launches[g].nruns), dim3(64), args, 0,
                                              stream_of(g)));
                }
            // bound the event pool on long calls: the oldest launches are long done
            if (h->pending.size() >= 512 && (rc = collect_timings(h, 256))) return rc;
            h->sweep_counter += n;
            h->phase_counter += n;
            c += n;
        }
        if (n_cycles > 0 && h->cfg.recompute_every > 0 && h->sweep_counter % h->cfg.recompute_every == 0) {
            for (size_t g = 0; g < ng; ++g) {
                const Launch &L = h->launches[g];
                SweepArgs a = base_args(h);
                a.tiles = h->d_tiles.p + L.tile0;
                a.sweep0 = h->sweep_counter;
                a.cycle0 = h->phase_counter;
                a.n_cycles = 0;
                a.init_run = -1;
                void *args[] = {&a};
                HIPCHK(h, hipLaunchKernel((const void *)L.fn, dim3(L.ntiles), dim3(L.wgs), args, L.lds, stream_of(g)));
            }
            h->recompute_applied = h->sweep_counter;
        }
        if (ng > 1)
            for (size_t g = 0; g < ng; ++g) {
                HIPCHK(h, hipEventRecord(h->join[g], stream_of(g)));
                HIPCHK(h, hipStreamWaitEvent(h->stream, h->join[g], 0));
            }
        done = n_cycles;
    }
    while (done < n_cycles) {
        const int64_t chunk = std::min<int64_t>(n_cycles - done, h->cfg.cycles_per_launch);
        SweepArgs a = base_args(h);
        a.sweep0 = h->sweep_counter;
        a.cycle0 = h->phase_counter;
        a.n_cycles = (int32_t)chunk;
        a.sampling = h->production;
        a.init_run = -1;
        if (h->trace_pending > 0) {
            if (chunk != n_cycles || chunk > h->trace_cycles)
                return fail(h, PTMC_E_STATE, "traced cycles must be run in one ptmc_cycles call of at most %d cycles",
                            h->trace_cycles);
            a.trace_de = h->d_trace_de.p;
            a.trace_acc = h->d_trace_acc.p;
            a.trace_idx = h->d_trace_idx.p;
            a.trace_cycles = h->trace_cycles;
            h->trace_pending = 0;
        }
        if ((rc = launch_all(h, a))) return rc;
        // bound the event pool on long calls
        if (h->pending.size() >= 512 && (rc = collect_timings(h, 256))) return rc;
        h->sweep_counter += chunk;
        h->phase_counter += chunk;
        done += chunk;
    }
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->kernel_ms += ms;
    // the per-launch timings are folded in when they are read
    // (ptmc_kernel_time, ptmc_group_info) or when many are pending
    return PTMC_OK;
}

int ptmc_sync(ptmc_handle h)
{
    if (!h) return PTMC_E_ARG;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->prepared && !h->runs.empty()) {
        std::vector<int32_t> e(h->runs.size());
        HIPCHK(h, hipMemcpy(e.data(), h->d_run_error.p, e.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (size_t r = 0; r < e.size(); ++r)
            if (e[r]) return fail(h, PTMC_E_SPHERE, "Error: too many moves out of binding sphere (run %zu)", r);
    }
    return PTMC_OK;
}

int ptmc_set_histogram(ptmc_handle h, int32_t run, double en_min, double en_max, double v_min, double v_max)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    RunHost &rh = h->runs[run];
    rh.info.en_min = en_min;
    rh.info.delta_en = (en_max - en_min) / (double)(h->cfg.n_bin - 1);
    rh.info.v_min = v_min;
    rh.info.delta_v = (v_max - v_min) / (double)h->cfg.n_bin;
    HIPCHK(h, hipMemcpyAsync(h->d_runs.p + run, &rh.info, sizeof(RunInfo), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_hist.p + (size_t)rh.info.temp_base * h->hist_stride, 0,
                             sizeof(uint32_t) * (size_t)rh.d.n_temps * h->hist_stride, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return PTMC_OK;
}

int ptmc_reset_counters(ptmc_handle h)
{
    if (!h) return PTMC_E_ARG;
    if (!h->prepared) return fail(h, PTMC_E_STATE, "engine not prepared");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipMemsetAsync(h->d_count_atom.p, 0, sizeof(long long) * 2 * h->n_temps, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_count_vol.p, 0, sizeof(long long) * 2 * h->n_temps, h->stream));
    if (h->d_count_volxy.p) {
        HIPCHK(h, hipMemsetAsync(h->d_count_volxy.p, 0, sizeof(long long) * 2 * h->n_temps, h->stream));
        HIPCHK(h, hipMemsetAsync(h->d_count_volz.p, 0, sizeof(long long) * 2 * h->n_temps, h->stream));
    }
    HIPCHK(h, hipMemsetAsync(h->d_count_exc.p, 0, sizeof(long long) * 2 * h->n_temps, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return PTMC_OK;
}

static int fetch_slots(ptmc_engine *h, int run, std::vector<int32_t> &temp, std::vector<double> &en,
                       std::vector<double> &box)
{
    const RunHost &rh = h->runs[run];
    const int R = rh.d.n_temps, s0 = rh.info.slot_base;
    temp.resize(R);
    en.resize(R);
    box.resize(R);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(temp.data(), h->d_slot_temp.p + s0, R * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(en.data(), h->d_slot_en.p + s0, R * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(box.data(), h->d_slot_box.p + s0, R * sizeof(double), hipMemcpyDeviceToHost));
    return PTMC_OK;
}

int ptmc_get_energies(ptmc_handle h, int32_t run, double *en)
{
    return ptmc_get_state(h, run, nullptr, en, nullptr);
}

int ptmc_get_state(ptmc_handle h, int32_t run, double *x, double *en, double *box)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    std::vector<int32_t> temp;
    std::vector<double> e, b;
    if ((rc = fetch_slots(h, run, temp, e, b))) return rc;
    const RunHost &rh = h->runs[run];
    const int R = rh.d.n_temps, N = rh.d.n_atoms;
    const bool has_h = h->cfg.boundary == BC_RECT || h->cfg.boundary == BC_RHOMB;
    std::vector<double> bh(R, 0.0);
    if (box && has_h)
        HIPCHK(h, hipMemcpy(bh.data(), h->d_slot_boxh.p + rh.info.slot_base, R * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < R; ++s) {
        const int t = temp[s];
        if (t < 0 || t >= R) return fail(h, PTMC_E_DEVICE, "corrupt temperature label");
        if (en) en[t] = e[s];
        if (box && has_h) {
            box[2 * t] = b[s];
            box[2 * t + 1] = bh[s];
        } else if (box) {
            box[t] = b[s];
        }
    }
    if (x && slot_f32(h)) {
        const int np = slot_np(h);
        std::vector<float> blk((size_t)R * 3 * np);
        HIPCHK(h, hipMemcpy(blk.data(), h->d_coords.p + (size_t)rh.info.slot_base * 3 * np * sizeof(float),
                            blk.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int s = 0; s < R; ++s)
            for (int j = 0; j < N; ++j)
                for (int c = 0; c < 3; ++c)
                    x[((size_t)temp[s] * N + j) * 3 + c] = blk[((size_t)s * 3 + c) * np + j];
    } else if (x && is_nnp2(h)) {
        std::vector<double> blk((size_t)R * 3 * NNP_NP);
        HIPCHK(h, hipMemcpy(blk.data(), h->d_coords.p + (size_t)rh.info.slot_base * 3 * NNP_NP * sizeof(double),
                            blk.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int s = 0; s < R; ++s)
            for (int j = 0; j < N; ++j)
                for (int c = 0; c < 3; ++c)
                    x[((size_t)temp[s] * N + j) * 3 + c] = blk[((size_t)s * 3 + c) * NNP_NP + j];
    } else if (x) {
      for (const Seg &g : rh.segs) {
        const TileInfo &ti = h->tiles[g.tile];
        const int A = ti.n_slab, kl = rh.kl;
        const size_t rsize = h->cfg.precision == 0 ? 8 : 4;
        const size_t nblk = (size_t)3 * A * rh.lanes;
        std::vector<unsigned char> blk(nblk * rsize);
        HIPCHK(h, hipMemcpy(blk.data(), h->d_coords.p + (size_t)ti.coord_off * rsize, blk.size(), hipMemcpyDeviceToHost));
        for (int s = g.t0; s < g.t0 + g.count; ++s) {
            const int t = temp[s], lg = g.local0 + (s - g.t0);
            for (int j = 0; j < N; ++j) {
                const int sl = j / kl, lane = lg * kl + j % kl;
                for (int c = 0; c < 3; ++c) {
                    const size_t e_ = ((size_t)sl * 3 + c) * rh.lanes + lane;
                    double v;
                    if (rsize == 8) memcpy(&v, &blk[e_ * 8], 8);
                    else {
                        float f;
                        memcpy(&f, &blk[e_ * 4], 4);
                        v = f;
                    }
                    x[((size_t)t * N + j) * 3 + c] = v;
                }
            }
        }
      }
    }
    return PTMC_OK;
}

int ptmc_get_counters(ptmc_handle h, int32_t run, double *max_displ, int64_t *count_atom, int64_t *count_vol,
                      int64_t *count_exc, double *ham)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t off = (size_t)2 * rh.info.temp_base, n = (size_t)2 * rh.d.n_temps;
    if (max_displ) HIPCHK(h, hipMemcpy(max_displ, h->d_displ.p + off, n * 8, hipMemcpyDeviceToHost));
    if (count_atom) HIPCHK(h, hipMemcpy(count_atom, h->d_count_atom.p + off, n * 8, hipMemcpyDeviceToHost));
    if (count_vol) HIPCHK(h, hipMemcpy(count_vol, h->d_count_vol.p + off, n * 8, hipMemcpyDeviceToHost));
    if (count_exc) HIPCHK(h, hipMemcpy(count_exc, h->d_count_exc.p + off, n * 8, hipMemcpyDeviceToHost));
    if (ham) HIPCHK(h, hipMemcpy(ham, h->d_ham.p + off, n * 8, hipMemcpyDeviceToHost));
    return PTMC_OK;
}

int ptmc_get_counters_sep(ptmc_handle h, int32_t run, double *max_displ_sep, int64_t *count_vol_xy,
                          int64_t *count_vol_z)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t off = (size_t)2 * rh.info.temp_base, n = (size_t)2 * rh.d.n_temps;
    if (max_displ_sep) HIPCHK(h, hipMemcpy(max_displ_sep, h->d_displ_sep.p + off, n * 8, hipMemcpyDeviceToHost));
    if (count_vol_xy) HIPCHK(h, hipMemcpy(count_vol_xy, h->d_count_volxy.p + off, n * 8, hipMemcpyDeviceToHost));
    if (count_vol_z) HIPCHK(h, hipMemcpy(count_vol_z, h->d_count_volz.p + off, n * 8, hipMemcpyDeviceToHost));
    return PTMC_OK;
}

int ptmc_set_counters_sep(ptmc_handle h, int32_t run, const double *max_displ_sep, const int64_t *count_vol_xy,
                          const int64_t *count_vol_z)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t off = (size_t)2 * rh.info.temp_base, n = (size_t)2 * rh.d.n_temps;
    if (max_displ_sep) HIPCHK(h, hipMemcpy(h->d_displ_sep.p + off, max_displ_sep, n * 8, hipMemcpyHostToDevice));
    if (count_vol_xy) HIPCHK(h, hipMemcpy(h->d_count_volxy.p + off, count_vol_xy, n * 8, hipMemcpyHostToDevice));
    if (count_vol_z) HIPCHK(h, hipMemcpy(h->d_count_volz.p + off, count_vol_z, n * 8, hipMemcpyHostToDevice));
    return PTMC_OK;
}

int ptmc_hist_stride(ptmc_handle h) { return h ? h->hist_stride : -1; }

int ptmc_get_histograms(ptmc_handle h, int32_t run, double *hist)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    if (!hist) return fail(h, PTMC_E_ARG, "null histogram buffer");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t n = (size_t)rh.d.n_temps * h->hist_stride;
    std::vector<uint32_t> tmp(n);
    HIPCHK(h, hipMemcpy(tmp.data(), h->d_hist.p + (size_t)rh.info.temp_base * h->hist_stride, n * 4,
                        hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) hist[i] = (double)tmp[i];
    return PTMC_OK;
}

int ptmc_set_histograms(ptmc_handle h, int32_t run, const double *hist)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    if (!hist) return fail(h, PTMC_E_ARG, "null histogram buffer");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t n = (size_t)rh.d.n_temps * h->hist_stride;
    std::vector<uint32_t> tmp(n);
    for (size_t i = 0; i < n; ++i) {
        if (!(hist[i] >= 0.0) || hist[i] > 4294967295.0 || hist[i] != std::floor(hist[i]))
            return fail(h, PTMC_E_ARG, "histogram entry %zu is not a count", i);
        tmp[i] = (uint32_t)hist[i];
    }
    HIPCHK(h, hipMemcpy(h->d_hist.p + (size_t)rh.info.temp_base * h->hist_stride, tmp.data(), n * 4,
                        hipMemcpyHostToDevice));
    return PTMC_OK;
}

int ptmc_seek(ptmc_handle h, int64_t cycle, int64_t phase_cycle)
{
    if (!h) return PTMC_E_ARG;
    if (cycle < 0 || phase_cycle < 0) return fail(h, PTMC_E_ARG, "negative cycle index");
    h->sweep_counter = cycle;
    h->recompute_applied = -1;
    h->phase_counter = phase_cycle;
    return PTMC_OK;
}

int ptmc_get_drift(ptmc_handle h, int32_t run, double *drift)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    std::vector<int32_t> temp;
    std::vector<double> e, b;
    if ((rc = fetch_slots(h, run, temp, e, b))) return rc;
    const RunHost &rh = h->runs[run];
    std::vector<double> dr(rh.d.n_temps);
    HIPCHK(h, hipMemcpy(dr.data(), h->d_slot_drift.p + rh.info.slot_base, dr.size() * 8, hipMemcpyDeviceToHost));
    for (int s = 0; s < rh.d.n_temps; ++s) drift[s] = dr[s];  // by slot
    return PTMC_OK;
}

int ptmc_set_state(ptmc_handle h, int32_t run, const double *x, const double *box, const double *max_displ,
                   const int64_t *count_atom, const int64_t *count_vol, const int64_t *count_exc, const double *ham)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const int R = rh.d.n_temps, N = rh.d.n_atoms;
    const size_t off = (size_t)2 * rh.info.temp_base, n = (size_t)2 * R;
    if (max_displ) HIPCHK(h, hipMemcpy(h->d_displ.p + off, max_displ, n * 8, hipMemcpyHostToDevice));
    if (count_atom) HIPCHK(h, hipMemcpy(h->d_count_atom.p + off, count_atom, n * 8, hipMemcpyHostToDevice));
    if (count_vol) HIPCHK(h, hipMemcpy(h->d_count_vol.p + off, count_vol, n * 8, hipMemcpyHostToDevice));
    if (count_exc) HIPCHK(h, hipMemcpy(h->d_count_exc.p + off, count_exc, n * 8, hipMemcpyHostToDevice));
    if (ham) HIPCHK(h, hipMemcpy(h->d_ham.p + off, ham, n * 8, hipMemcpyHostToDevice));
    if (x || box) {
        // slot s <- temperature s; positions rewritten, energies recomputed
        std::vector<int32_t> temp(R);
        for (int s = 0; s < R; ++s) temp[s] = s;
        HIPCHK(h, hipMemcpy(h->d_slot_temp.p + rh.info.slot_base, temp.data(), R * 4, hipMemcpyHostToDevice));
        if (box) {
            // lengths, or (length, height) pairs for rectangular and rhombic boxes;
            // a cubic box's height is its length
            const bool has_h = h->cfg.boundary == BC_RECT || h->cfg.boundary == BC_RHOMB;
            std::vector<double> bl(R), bh(R);
            for (int t = 0; t < R; ++t) {
                bl[t] = has_h ? box[2 * t] : box[t];
                bh[t] = has_h ? box[2 * t + 1] : box[t];
                if (h->cfg.boundary != BC_SPHERE && !(bl[t] > 0 && bh[t] > 0))
                    return fail(h, PTMC_E_ARG, "box dimensions must be positive");
            }
            HIPCHK(h, hipMemcpy(h->d_slot_box.p + rh.info.slot_base, bl.data(), R * 8, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->d_slot_boxh.p + rh.info.slot_base, bh.data(), R * 8, hipMemcpyHostToDevice));
        }
        if (is_slot(h) || h->spread) {
            std::vector<int32_t> ts(R);
            for (int s = 0; s < R; ++s) ts[s] = rh.info.slot_base + s;
            HIPCHK(h, hipMemcpy(h->d_temp_slot.p + rh.info.temp_base, ts.data(), R * 4, hipMemcpyHostToDevice));
        }
        if (x && slot_f32(h)) {
            const int np = slot_np(h);
            std::vector<float> blk((size_t)R * 3 * np, 0.0f);
            for (int s = 0; s < R; ++s)
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c)
                        blk[((size_t)s * 3 + c) * np + j] = (float)x[((size_t)s * N + j) * 3 + c];
            HIPCHK(h, hipMemcpy(h->d_coords.p + (size_t)rh.info.slot_base * 3 * np * sizeof(float), blk.data(),
                                blk.size() * sizeof(float), hipMemcpyHostToDevice));
        } else if (x && is_nnp2(h)) {
            std::vector<double> blk((size_t)R * 3 * NNP_NP, 0.0);
            for (int s = 0; s < R; ++s)
                for (int j = 0; j < N; ++j)
                    for (int c = 0; c < 3; ++c) blk[((size_t)s * 3 + c) * NNP_NP + j] = x[((size_t)s * N + j) * 3 + c];
            HIPCHK(h, hipMemcpy(h->d_coords.p + (size_t)rh.info.slot_base * 3 * NNP_NP * sizeof(double), blk.data(),
                                blk.size() * sizeof(double), hipMemcpyHostToDevice));
        } else if (x) {
          for (const Seg &g : rh.segs) {
            const TileInfo &ti = h->tiles[g.tile];
            const int A = ti.n_slab, kl = rh.kl;
            const size_t rsize = h->cfg.precision == 0 ? 8 : 4;
            const size_t nblk = (size_t)3 * A * rh.lanes;
            std::vector<unsigned char> blk(nblk * rsize);
            HIPCHK(h, hipMemcpy(blk.data(), h->d_coords.p + (size_t)ti.coord_off * rsize, blk.size(), hipMemcpyDeviceToHost));
            for (int s = g.t0; s < g.t0 + g.count; ++s) {
                const int lg = g.local0 + (s - g.t0);
                for (int j = 0; j < N; ++j) {
                    const int sl = j / kl, lane = lg * kl + j % kl;
                    for (int c = 0; c < 3; ++c) {
                        const size_t e_ = ((size_t)sl * 3 + c) * rh.lanes + lane;
                        const double v = x[((size_t)s * N + j) * 3 + c];
                        if (rsize == 8) memcpy(&blk[e_ * 8], &v, 8);
                        else {
                            float f = (float)v;
                            memcpy(&blk[e_ * 4], &f, 4);
                        }
                    }
                }
            }
            HIPCHK(h, hipMemcpy(h->d_coords.p + (size_t)ti.coord_off * rsize, blk.data(), blk.size(), hipMemcpyHostToDevice));
          }
        }
        SweepArgs a = base_args(h);
        a.init_run = run;
        if ((rc = launch_all(h, a, true))) return rc;
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return PTMC_OK;
}

int ptmc_set_trace(ptmc_handle h, int32_t n_cycles)
{
    if (!h) return PTMC_E_ARG;
    int rc = ptmc_prepare(h);
    if (rc) return rc;
    if (n_cycles <= 0) return fail(h, PTMC_E_ARG, "trace length must be positive");
    int maxN = 0;
    for (auto &rh : h->runs) maxN = std::max(maxN, rh.d.n_atoms);
    const int nsteps = h->cfg.ensemble == ENS_NPT ? maxN + 1 : (is_nnp2(h) ? maxN + h->nnp2.n_swaps : maxN);
    const size_t n = (size_t)h->n_temps * n_cycles * nsteps;
    for (auto &rh : h->runs)
        if (rh.d.n_atoms != maxN) return fail(h, PTMC_E_UNSUPPORTED, "trace needs equal atom counts");
    if ((rc = dalloc(h, h->d_trace_de, n))) return rc;
    if ((rc = dalloc(h, h->d_trace_acc, n))) return rc;
    if ((rc = dalloc(h, h->d_trace_idx, n))) return rc;
    if ((rc = dalloc(h, h->d_trace_aux, n))) return rc;
    // NaN (all bits set) where no surrogate change is recorded
    HIPCHK(h, hipMemset(h->d_trace_aux.p, 0xff, n * sizeof(double)));
    h->trace_cycles = n_cycles;
    h->trace_pending = 1;
    h->trace_nsteps = nsteps;
    return PTMC_OK;
}

int ptmc_get_trace(ptmc_handle h, int32_t run, double *de, int8_t *acc, int32_t *idx)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    if (!h->d_trace_de.p) return fail(h, PTMC_E_STATE, "no trace recorded");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t per = (size_t)h->trace_cycles * h->trace_nsteps;
    const size_t off = (size_t)rh.info.temp_base * per, n = (size_t)rh.d.n_temps * per;
    if (de) HIPCHK(h, hipMemcpy(de, h->d_trace_de.p + off, n * 8, hipMemcpyDeviceToHost));
    if (acc) HIPCHK(h, hipMemcpy(acc, h->d_trace_acc.p + off, n, hipMemcpyDeviceToHost));
    if (idx) HIPCHK(h, hipMemcpy(idx, h->d_trace_idx.p + off, n * 4, hipMemcpyDeviceToHost));
    return PTMC_OK;
}

int ptmc_get_trace_aux(ptmc_handle h, int32_t run, double *aux)
{
    int rc = check_run(h, run);
    if (rc) return rc;
    if (!h->d_trace_aux.p) return fail(h, PTMC_E_STATE, "no trace recorded");
    if (!aux) return fail(h, PTMC_E_ARG, "null buffer");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const RunHost &rh = h->runs[run];
    const size_t per = (size_t)h->trace_cycles * h->trace_nsteps;
    const size_t off = (size_t)rh.info.temp_base * per, n = (size_t)rh.d.n_temps * per;
    HIPCHK(h, hipMemcpy(aux, h->d_trace_aux.p + off, n * 8, hipMemcpyDeviceToHost));
    return PTMC_OK;
}

int ptmc_plan_info(ptmc_handle h, int32_t *n_tiles, int32_t *n_launch_groups, int64_t *n_lanes)
{
    if (!h) return PTMC_E_ARG;
    if (!h->prepared) return fail(h, PTMC_E_STATE, "engine not prepared");
    if (n_tiles) *n_tiles = is_slot(h) ? h->n_slots : (int32_t)h->tiles.size();
    if (n_launch_groups) *n_launch_groups = (int32_t)h->launches.size();
    if (n_lanes) {
        int64_t s = 0;
        for (auto &L : h->launches) s += (int64_t)L.ntiles * L.wgs;
        *n_lanes = s;
    }
    return PTMC_OK;
}

int ptmc_run_lanes(ptmc_handle h, int32_t run)
{
    if (!h || run < 0 || run >= (int)h->runs.size() || !h->prepared) return -1;
    return h->runs[run].kl;
}

int ptmc_kernel_time(ptmc_handle h, double *total_ms, int64_t *n_launches, int32_t reset)
{
    if (!h) return PTMC_E_ARG;
    if (!h->pending.empty()) {
        HIPCHK(h, hipSetDevice(h->cfg.device));
        HIPCHK(h, hipDeviceSynchronize());
        int rc = collect_timings(h);
        if (rc) return rc;
    }
    if (total_ms) *total_ms = h->kernel_ms;
    if (n_launches) *n_launches = h->n_launch;
    if (reset) {
        h->kernel_ms = 0.0;
        h->n_launch = 0;
        for (auto &L : h->launches) {
            L.ms = 0.0;
            L.n_done = 0;
        }
    }
    return PTMC_OK;
}

int ptmc_group_info(ptmc_handle h, int32_t group, int32_t *lanes_per_replica, int32_t *workgroup_lanes,
                    int32_t *n_tiles, double *total_ms, int64_t *n_launches)
{
    if (!h) return PTMC_E_ARG;
    if (!h->prepared) return fail(h, PTMC_E_STATE, "engine not prepared");
    if (group < 0 || group >= (int)h->launches.size()) return fail(h, PTMC_E_ARG, "launch group %d out of range", group);
    if (!h->pending.empty()) {
        HIPCHK(h, hipSetDevice(h->cfg.device));
        HIPCHK(h, hipDeviceSynchronize());
        int rc = collect_timings(h);
        if (rc) return rc;
    }
    const Launch &L = h->launches[group];
    if (lanes_per_replica) *lanes_per_replica = L.kl;
    if (workgroup_lanes) *workgroup_lanes = L.wgs;
    if (n_tiles) *n_tiles = L.ntiles;
    if (total_ms) *total_ms = L.ms;
    if (n_launches) *n_launches = L.n_done;
    return PTMC_OK;
}

int ptmc_run_group(ptmc_handle h, int32_t run)
{
    if (!h || run < 0 || run >= (int)h->runs.size() || !h->prepared) return -1;
    if (is_slot(h)) return 0;
    const int t = h->runs[run].tile;
    for (size_t k = 0; k < h->launches.size(); ++k)
        if (t >= h->launches[k].tile0 && t < h->launches[k].tile0 + h->launches[k].ntiles) return (int)k;
    return -1;
}


// ------------------------------------------------------------------ surrogate

// the delayed-acceptance sweep kernel of the engine's potential
static SweepFn da_sweep_fn(ptmc_engine *h)
{
    if (!is_pair_slot(h)) return nnp_sweep_da_fn();
    SweepFn fs = nullptr, fi = nullptr;
    pair_slot_fns(h->pk, h->kc, true, &fs, &fi);
    return fs;
}

static int sur_check(ptmc_engine *h)
{
    if (!h) return PTMC_E_ARG;
    if (!h->sur.init) return fail(h, PTMC_E_STATE, "surrogate not initialised (ptmc_surrogate_init)");
    return PTMC_OK;
}

int ptmc_surrogate_init(ptmc_handle h, int32_t capacity, uint64_t seed)
{
    if (!h) return PTMC_E_ARG;
    int rc = ptmc_prepare(h);
    if (rc) return rc;
    if (!is_nnp(h) && !is_pair_slot(h))
        return fail(h, PTMC_E_UNSUPPORTED, "the surrogate filters the network potential or a pair potential in pair "
                    "slots (ptmc_config.pair_slots)");
    if (h->nnp_np != NNP_NP)
        return fail(h, PTMC_E_UNSUPPORTED, "the surrogate covers configurations of at most %d atoms", NNP_NP);
    if (capacity < 1) return fail(h, PTMC_E_ARG, "sample capacity must be positive");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int dev = 0, max_lds = 0;
    HIPCHK(h, hipGetDevice(&dev));
    HIPCHK(h, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t lds = is_pair_slot(h) ? pair_slot_lds_bytes(true) : h->nnp_lds + sur_lds_floats() * sizeof(float);
    if ((long long)lds > max_lds) return fail(h, PTMC_E_UNSUPPORTED, "delayed acceptance needs %zu B of LDS", lds);
    HIPCHK(h, hipFuncSetAttribute((const void *)da_sweep_fn(h), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    h->sur.cap = capacity;
    h->sur.seed = seed;
    h->sur.n_samples = 0;
    h->sur.t = 0;
    if ((rc = dalloc(h, h->d_sur_D, (size_t)capacity * NNP_NP * SUR_K))) return rc;
    if ((rc = dalloc(h, h->d_sur_E, (size_t)capacity))) return rc;
    if ((rc = dalloc(h, h->d_sur_N, (size_t)capacity))) return rc;
    if ((rc = dalloc(h, h->d_sur_pred, (size_t)capacity))) return rc;
    if ((rc = dalloc(h, h->d_sur_idx, (size_t)capacity))) return rc;
    if ((rc = dalloc(h, h->d_sur_m, SUR_P))) return rc;
    if ((rc = dalloc(h, h->d_sur_v, SUR_P))) return rc;
    if ((rc = dalloc(h, h->d_sur_grad, SUR_P))) return rc;
    if ((rc = dalloc(h, h->d_da_stats, (size_t)3 * h->n_slots))) return rc;
    // He-style start, b3 = mean atomic energy of the current states
    std::vector<double> en(h->n_slots);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(en.data(), h->d_slot_en.p, en.size() * 8, hipMemcpyDeviceToHost));
    double e_atom = 0.0;
    for (auto &rh : h->runs)
        for (int t = 0; t < rh.d.n_temps; ++t) e_atom += en[rh.info.slot_base + t] / rh.d.n_atoms;
    e_atom /= h->n_slots;
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(SUR_P, 0.0f);
    int o = 0;
    for (int e = 0; e < SUR_K * SUR_H; ++e) w[o++] = (float)(nd(gen) * std::sqrt(2.0 / SUR_K));
    o += SUR_H;
    for (int e = 0; e < SUR_H * SUR_H; ++e) w[o++] = (float)(nd(gen) * std::sqrt(2.0 / SUR_H));
    o += SUR_H;
    for (int e = 0; e < SUR_H; ++e) w[o++] = (float)(nd(gen) * std::sqrt(1.0 / SUR_H));
    w[o] = (float)e_atom;
    if ((rc = upload(h, h->d_sur_w, w))) return rc;
    h->sur.init = true;
    return PTMC_OK;
}

int ptmc_surrogate_collect(ptmc_handle h, int32_t every)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (every < 0) return fail(h, PTMC_E_ARG, "negative collection period");
    h->sur.collect_every = every;
    return PTMC_OK;
}

int ptmc_surrogate_samples(ptmc_handle h, int64_t *n)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (n) *n = std::min<long long>(h->sur.n_samples, h->sur.cap);
    return PTMC_OK;
}

int ptmc_surrogate_enable(ptmc_handle h, int32_t on)
{
    int rc = sur_check(h);
    if (rc) return rc;
    h->sur.da = on != 0;
    Launch &L = h->launches[0];
    if (is_pair_slot(h)) {
        SweepFn fs = nullptr, fi = nullptr;
        pair_slot_fns(h->pk, h->kc, h->sur.da, &fs, &fi);
        L.fn = fs;
        L.lds = pair_slot_lds_bytes(h->sur.da);
        return PTMC_OK;
    }
    L.fn = h->sur.da ? nnp_sweep_da_fn() : nnp_sweep_fn(h->nnp_np);
    L.lds = h->nnp_lds + (h->sur.da ? sur_lds_floats() * sizeof(float) : 0);
    return PTMC_OK;
}

int ptmc_surrogate_get_params(ptmc_handle h, double *p)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (!p) return fail(h, PTMC_E_ARG, "null buffer");
    std::vector<float> w(SUR_P);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(w.data(), h->d_sur_w.p, SUR_P * 4, hipMemcpyDeviceToHost));
    for (int e = 0; e < SUR_P; ++e) p[e] = w[e];
    return PTMC_OK;
}

int ptmc_surrogate_set_params(ptmc_handle h, const double *p)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (!p) return fail(h, PTMC_E_ARG, "null buffer");
    std::vector<float> w(SUR_P);
    for (int e = 0; e < SUR_P; ++e) w[e] = (float)p[e];
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(h->d_sur_w.p, w.data(), SUR_P * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(h->d_sur_m.p, 0, SUR_P * 4));
    HIPCHK(h, hipMemset(h->d_sur_v.p, 0, SUR_P * 4));
    h->sur.t = 0;
    return PTMC_OK;
}

// descriptors of n host configurations (n x n_atoms x 3) into d_sur_x scratch
static int sur_upload_configs(ptmc_engine *h, int32_t n, int32_t n_atoms, const double *x, DevBuf<float> &D,
                              DevBuf<int32_t> &Nb)
{
    if (n < 1 || n_atoms < 2 || n_atoms > NNP_NP || !x) return fail(h, PTMC_E_ARG, "bad configurations");
    std::vector<float> buf((size_t)n * 3 * NNP_NP, 0.0f);
    for (int s = 0; s < n; ++s)
        for (int j = 0; j < n_atoms; ++j)
            for (int c = 0; c < 3; ++c) buf[((size_t)s * 3 + c) * NNP_NP + j] = (float)x[((size_t)s * n_atoms + j) * 3 + c];
    int rc;
    if ((rc = upload(h, h->d_sur_x, buf))) return rc;
    std::vector<int32_t> nv(n, n_atoms);
    if ((rc = upload(h, Nb, nv))) return rc;
    if ((rc = dalloc(h, D, (size_t)n * NNP_NP * SUR_K))) return rc;
    HIPCHK(h, sur_launch_desc(h->d_sur_x.p, Nb.p, D.p, (float)sur_cutoff_of(h), n, h->stream));
    return PTMC_OK;
}

int ptmc_surrogate_predict(ptmc_handle h, int32_t n, int32_t n_atoms, const double *x, double *e_hat)
{
    int rc = sur_check(h);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    DevBuf<float> D;
    DevBuf<int32_t> Nb, idx;
    DevBuf<double> E, pred;
    if ((rc = sur_upload_configs(h, n, n_atoms, x, D, Nb))) return rc;
    std::vector<int32_t> iv(n);
    for (int s = 0; s < n; ++s) iv[s] = s;
    if ((rc = upload(h, idx, iv)) || (rc = dalloc(h, E, n)) || (rc = dalloc(h, pred, n))) return rc;
    HIPCHK(h, sur_launch_train(D.p, E.p, Nb.p, idx.p, h->d_sur_w.p, nullptr, pred.p, 0, n, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(e_hat, pred.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    D.free_(); Nb.free_(); idx.free_(); E.free_(); pred.free_();
    return PTMC_OK;
}

int ptmc_surrogate_grad(ptmc_handle h, int32_t n, int32_t n_atoms, const double *x, const double *e, double *loss,
                        double *grad)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (!e) return fail(h, PTMC_E_ARG, "null energies");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    DevBuf<float> D, part;
    DevBuf<int32_t> Nb, idx;
    DevBuf<double> E, pred;
    if ((rc = sur_upload_configs(h, n, n_atoms, x, D, Nb))) return rc;
    std::vector<int32_t> iv(n);
    for (int s = 0; s < n; ++s) iv[s] = s;
    std::vector<double> ev(e, e + n);
    if ((rc = upload(h, idx, iv)) || (rc = upload(h, E, ev)) || (rc = dalloc(h, pred, n)) ||
        (rc = dalloc(h, part, (size_t)n * SUR_P)))
        return rc;
    HIPCHK(h, sur_launch_train(D.p, E.p, Nb.p, idx.p, h->d_sur_w.p, part.p, pred.p, 1, n, h->stream));
    HIPCHK(h, sur_launch_reduce(part.p, n, h->d_sur_grad.p, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<double> pv(n);
    HIPCHK(h, hipMemcpy(pv.data(), pred.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (loss) {
        double l = 0.0;
        for (int s = 0; s < n; ++s) {
            const double r = (pv[s] - e[s]) / n_atoms;
            l += r * r;
        }
        *loss = l / n;
    }
    if (grad) {
        std::vector<float> g(SUR_P);
        HIPCHK(h, hipMemcpy(g.data(), h->d_sur_grad.p, SUR_P * 4, hipMemcpyDeviceToHost));
        for (int p = 0; p < SUR_P; ++p) grad[p] = g[p];
    }
    D.free_(); part.free_(); Nb.free_(); idx.free_(); E.free_(); pred.free_();
    return PTMC_OK;
}

// RMSE (Hartree per configuration) of the samples listed in idx[0..n)
static int sur_rmse(ptmc_engine *h, const std::vector<int32_t> &iv, const std::vector<double> &Eall, double *out)
{
    if (iv.empty()) {
        *out = 0.0;
        return PTMC_OK;
    }
    HIPCHK(h, hipMemcpy(h->d_sur_idx.p, iv.data(), iv.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(h, sur_launch_train(h->d_sur_D.p, h->d_sur_E.p, h->d_sur_N.p, h->d_sur_idx.p, h->d_sur_w.p, nullptr,
                               h->d_sur_pred.p, 0, (int)iv.size(), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<double> pv(iv.size());
    HIPCHK(h, hipMemcpy(pv.data(), h->d_sur_pred.p, pv.size() * 8, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (size_t k = 0; k < iv.size(); ++k) {
        const double r = pv[k] - Eall[iv[k]];
        s += r * r;
    }
    *out = std::sqrt(s / iv.size());
    return PTMC_OK;
}

int ptmc_surrogate_train(ptmc_handle h, int32_t epochs, int32_t batch, double lr, double valid_fraction,
                         double *rmse_train, double *rmse_valid)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (epochs < 0 || batch < 1 || !(lr > 0) || !(valid_fraction >= 0 && valid_fraction < 1))
        return fail(h, PTMC_E_ARG, "bad training settings");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int n = (int)std::min<long long>(h->sur.n_samples, h->sur.cap);
    if (n < 2) return fail(h, PTMC_E_STATE, "fewer than 2 samples collected");
    std::vector<double> Eall(h->sur.cap);
    HIPCHK(h, hipMemcpy(Eall.data(), h->d_sur_E.p, Eall.size() * 8, hipMemcpyDeviceToHost));
    // fixed split: a seeded permutation, the last valid_fraction held out
    std::vector<int32_t> perm(n);
    for (int s = 0; s < n; ++s) perm[s] = s;
    std::mt19937_64 gen(h->sur.seed ^ 0x5eed5eedull);
    std::shuffle(perm.begin(), perm.end(), gen);
    const int n_valid = (int)(valid_fraction * n);
    const int n_train = n - n_valid;
    std::vector<int32_t> tr(perm.begin(), perm.begin() + n_train), va(perm.begin() + n_train, perm.end());
    const int B = std::min(batch, n_train);
    if ((size_t)B * SUR_P > h->d_sur_partial.n) {
        if ((rc = dalloc(h, h->d_sur_partial, (size_t)B * SUR_P))) return rc;
    }
    std::mt19937_64 gen2(h->sur.seed + (uint64_t)h->sur.t);
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(tr.begin(), tr.end(), gen2);
        HIPCHK(h, hipMemcpy(h->d_sur_idx.p, tr.data(), tr.size() * 4, hipMemcpyHostToDevice));
        for (int b0 = 0; b0 + B <= n_train; b0 += B) {
            HIPCHK(h, sur_launch_train(h->d_sur_D.p, h->d_sur_E.p, h->d_sur_N.p, h->d_sur_idx.p + b0, h->d_sur_w.p,
                                       h->d_sur_partial.p, h->d_sur_pred.p, 1, B, h->stream));
            HIPCHK(h, sur_launch_reduce(h->d_sur_partial.p, B, h->d_sur_grad.p, h->stream));
            h->sur.t += 1;
            HIPCHK(h, sur_launch_adam(h->d_sur_w.p, h->d_sur_grad.p, h->d_sur_m.p, h->d_sur_v.p, h->sur.t, (float)lr,
                                      h->stream));
        }
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (tr.size() > 4096) tr.resize(4096);
    if (rmse_train && (rc = sur_rmse(h, tr, Eall, rmse_train))) return rc;
    if (rmse_valid && (rc = sur_rmse(h, va, Eall, rmse_valid))) return rc;
    return PTMC_OK;
}

int ptmc_surrogate_stats(ptmc_handle h, int32_t run, int64_t *proposals, int64_t *stage1, int64_t *accepted)
{
    int rc = sur_check(h);
    if (rc) return rc;
    if (run < 0 || run >= (int)h->runs.size()) return fail(h, PTMC_E_ARG, "run index out of range");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<long long> st((size_t)3 * h->n_slots);
    HIPCHK(h, hipMemcpy(st.data(), h->d_da_stats.p, st.size() * 8, hipMemcpyDeviceToHost));
    const RunHost &rh = h->runs[run];
    long long p = 0, s1 = 0, acc = 0;
    for (int t = 0; t < rh.d.n_temps; ++t) {
        const int sl = rh.info.slot_base + t;
        p += st[3 * sl];
        s1 += st[3 * sl + 1];
        acc += st[3 * sl + 2];
    }
    if (proposals) *proposals = p;
    if (stage1) *stage1 = s1;
    if (accepted) *accepted = acc;
    return PTMC_OK;
}

}  // extern "C"

// ---- potential plugin functions: set_variables / initialise_energy /
// energy_update! on caller-supplied configurations (ptmc_plugin.hpp)

namespace {

// values of the potential's per-configuration state (the reference's
// PotentialVariables that these functions pass on): EAM components, network
// features; pair potentials keep none
int64_t plugin_nvars(const ptmc_engine *h, int n_atoms)
{
    if (h->pot.kind == POT_EAM) return 2 * (int64_t)n_atoms;
    if (h->pot.kind == POT_NNP) return (int64_t)n_atoms * h->nnp.nf;
    return 0;
}

int plugin_job(ptmc_engine *h, int32_t n_conf, int32_t n_atoms, const double *x, const double *d2, const double *box,
               bool update, PluginJob &j)
{
    if (!h->pot_set) return fail(h, PTMC_E_STATE, "no potential set");
    if (n_conf < 1 || n_atoms < 2 || n_atoms > PLUGIN_MAXN)
        return fail(h, PTMC_E_ARG, "need n_conf >= 1 and 2 <= n_atoms <= %d", PLUGIN_MAXN);
    const PotParams &p = h->pot;
    j.n_conf = n_conf;
    j.N = n_atoms;
    j.kind = p.kind;
    j.bc = h->cfg.boundary;
    j.div = h->cfg.pair_form == 1 ? 1 : 0;
    j.pot = p;
    const bool npt = (h->cfg.ensemble == 1 || h->cfg.ensemble == 2) && j.bc != BC_SPHERE;
    const bool dist_only = p.kind == POT_ELJ_EVEN || p.kind == POT_ELJ || p.kind == POT_EAM;
    if (d2 && !dist_only)
        return fail(h, PTMC_E_ARG, "distance matrices serve the distance-only potentials (ELJ, EAM); this one needs positions");
    if (!x && (update || !d2)) return fail(h, PTMC_E_ARG, "positions required");
    if (j.bc != BC_SPHERE && !box) return fail(h, PTMC_E_ARG, "box lengths required in a periodic box");
    switch (p.kind) {
    case POT_ELJ_EVEN:
    case POT_ELJ:
        if (p.ncoeff > 8) return fail(h, PTMC_E_UNSUPPORTED, "at most 8 ELJ coefficients");
        if (npt && p.kind == POT_ELJ)
            return fail(h, PTMC_E_UNSUPPORTED, "ELJPotential (odd powers) has no tail correction for NPT");
        j.pk = p.kind == POT_ELJ_EVEN ? (j.div ? PK_EVEN_R : PK_EVEN_H) : (j.div ? PK_ODD_R : PK_ODD_H);
        j.cutoff = npt ? 1 : 0;
        break;
    case POT_ELJB:
        j.pk = PK_ELJB;
        j.cutoff = npt ? 1 : 0;
        break;
    case POT_LUT: {
        j.pk = PK_LUT;
        j.cutoff = npt ? 1 : 0;
        if (!h->d_lut64.p) {
            const int rc = upload(h, h->d_lut64, h->lut_host);
            if (rc) return rc;
        }
        j.pot.lut64 = h->d_lut64.p;
        break;
    }
    case POT_EAM: break;
    case POT_NNP: {
        const auto &n = h->nnp;
        if (j.bc != BC_SPHERE)
            return fail(h, PTMC_E_UNSUPPORTED, "the network potential's plugin functions run in a spherical container");
        if (n.h1 > PLUGIN_MAXH || n.h2 > PLUGIN_MAXH) return fail(h, PTMC_E_UNSUPPORTED, "hidden layers wider than %d", PLUGIN_MAXH);
        if (!h->plug_nnp_up) {
            int rc = upload(h, h->d_plug_rad, n.rad64.empty() ? std::vector<double>(4, 0.0) : n.rad64);
            if (!rc) rc = upload(h, h->d_plug_ang, n.ang64.empty() ? std::vector<double>(6, 0.0) : n.ang64);
            if (!rc) rc = upload(h, h->d_plug_w, n.w64);
            if (rc) return rc;
            h->plug_nnp_up = true;
        }
        j.n_rad = n.n_rad;
        j.n_ang = n.n_ang;
        j.n_ang_built = n.q10 ? std::max(0, n.n_ang - n.n_rad) : n.n_ang;
        j.h1 = n.h1;
        j.h2 = n.h2;
        j.r_cut = n.r_cut;
        j.hardcore2 = n.hardcore2;
        j.rad = h->d_plug_rad.p;
        j.ang = h->d_plug_ang.p;
        j.w = h->d_plug_w.p;
        break;
    }
    default:
        return fail(h, PTMC_E_UNSUPPORTED, "no plugin functions for this potential (two species: ptmc_nnp2_energy)");
    }
    return PTMC_OK;
}

// device arena of one call: regions in doubles, 256-byte aligned
struct PlugLayout {
    size_t x = 0, d2 = 0, box = 0, index = 0, trial = 0, en_in = 0, vars_in = 0, en = 0, en_atom = 0, vars_out = 0,
           ws = 0, total = 0;
};

PlugLayout plug_layout(const PluginJob &j, int64_t nv, bool has_d2, bool has_box, bool update)
{
    const size_t C = (size_t)j.n_conf, N = (size_t)j.N;
    PlugLayout L;
    size_t o = 0;
    auto take = [&](size_t n) {
        const size_t at = o;
        o += (n + 31) / 32 * 32;
        return at;
    };
    L.x = take(C * N * 3);
    L.d2 = take(has_d2 ? C * N * N : 0);
    L.box = take(has_box ? C * 2 : 0);
    L.index = take(update ? (C + 1) / 2 : 0);
    L.trial = take(update ? C * 3 : 0);
    L.en_in = take(update ? C : 0);
    L.vars_in = take(update ? C * (size_t)nv : 0);
    L.en = take(C);
    L.en_atom = take(C * N);
    L.vars_out = take(C * (size_t)nv);
    L.ws = take(plugin_workspace_doubles(j));
    L.total = o;
    return L;
}

int plug_arena(ptmc_engine *h, size_t n)
{
    if (h->d_plug.n >= n) return PTMC_OK;
    h->d_plug.free_();
    HIPCHK(h, hipMalloc((void **)&h->d_plug.p, n * sizeof(double)));
    h->d_plug.n = n;
    return PTMC_OK;
}

int plug_put(ptmc_engine *h, size_t at, const void *src, size_t bytes)
{
    if (src && bytes) HIPCHK(h, hipMemcpyAsync(h->d_plug.p + at, src, bytes, hipMemcpyHostToDevice, h->stream));
    return PTMC_OK;
}

int plug_get(ptmc_engine *h, void *dst, size_t at, size_t bytes)
{
    if (dst && bytes) HIPCHK(h, hipMemcpyAsync(dst, h->d_plug.p + at, bytes, hipMemcpyDeviceToHost, h->stream));
    return PTMC_OK;
}

}  // namespace

extern "C" {

int ptmc_potential_vars(ptmc_handle h, int32_t n_atoms, int64_t *n_vars)
{
    if (!h || !n_vars || n_atoms < 1) return h ? fail(h, PTMC_E_ARG, "bad arguments") : PTMC_E_ARG;
    if (!h->pot_set) return fail(h, PTMC_E_STATE, "no potential set");
    *n_vars = plugin_nvars(h, n_atoms);
    return PTMC_OK;
}

int ptmc_initialise_energy(ptmc_handle h, int32_t n_conf, int32_t n_atoms, const double *x, const double *d2,
                           const double *box, double *en_tot, double *en_atom, double *vars)
{
    if (!h) return PTMC_E_ARG;
    if (!en_tot) return fail(h, PTMC_E_ARG, "en_tot required");
    PluginJob j;
    int rc = plugin_job(h, n_conf, n_atoms, x, d2, box, false, j);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const int64_t nv = plugin_nvars(h, n_atoms);
    const size_t C = (size_t)n_conf, N = (size_t)n_atoms;
    const PlugLayout L = plug_layout(j, nv, d2 != nullptr, box != nullptr, false);
    if ((rc = plug_arena(h, L.total))) return rc;
    double *A = h->d_plug.p;
    if ((rc = plug_put(h, L.x, x, C * N * 3 * 8)) || (rc = plug_put(h, L.d2, d2, C * N * N * 8)) ||
        (rc = plug_put(h, L.box, box, C * 2 * 8)))
        return rc;
    j.x = x ? A + L.x : nullptr;
    j.d2_in = d2 ? A + L.d2 : nullptr;
    j.box = box ? A + L.box : nullptr;
    j.en = A + L.en;
    j.en_atom = A + L.en_atom;
    j.vars_out = nv ? A + L.vars_out : nullptr;
    j.ws = A + L.ws;
    HIPCHK(h, plugin_initialise(j, h->stream));
    if ((rc = plug_get(h, en_tot, L.en, C * 8)) || (rc = plug_get(h, en_atom, L.en_atom, C * N * 8)) ||
        (rc = plug_get(h, nv ? vars : nullptr, L.vars_out, C * (size_t)nv * 8)))
        return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return PTMC_OK;
}

int ptmc_energy_update(ptmc_handle h, int32_t n_conf, int32_t n_atoms, const double *x, const double *d2,
                       const double *box, const int32_t *index, const double *trial, const double *en_tot,
                       const double *vars, double *new_en, double *new_en_atom, double *new_vars)
{
    if (!h) return PTMC_E_ARG;
    if (!index || !trial || !en_tot || !new_en) return fail(h, PTMC_E_ARG, "index, trial, en_tot and new_en required");
    PluginJob j;
    int rc = plugin_job(h, n_conf, n_atoms, x, d2, box, true, j);
    if (rc) return rc;
    const int64_t nv = plugin_nvars(h, n_atoms);
    if (nv && !vars) return fail(h, PTMC_E_ARG, "this potential's update needs its variables (ptmc_initialise_energy)");
    for (int c = 0; c < n_conf; ++c)
        if (index[c] < 0 || index[c] >= n_atoms) return fail(h, PTMC_E_ARG, "atom index %d out of range", index[c]);
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const size_t C = (size_t)n_conf, N = (size_t)n_atoms;
    const PlugLayout L = plug_layout(j, nv, d2 != nullptr, box != nullptr, true);
    if ((rc = plug_arena(h, L.total))) return rc;
    double *A = h->d_plug.p;
    if ((rc = plug_put(h, L.x, x, C * N * 3 * 8)) || (rc = plug_put(h, L.d2, d2, C * N * N * 8)) ||
        (rc = plug_put(h, L.box, box, C * 2 * 8)) || (rc = plug_put(h, L.index, index, C * 4)) ||
        (rc = plug_put(h, L.trial, trial, C * 3 * 8)) || (rc = plug_put(h, L.en_in, en_tot, C * 8)) ||
        (rc = plug_put(h, L.vars_in, nv ? vars : nullptr, C * (size_t)nv * 8)))
        return rc;
    // a network configuration that fails the hard-core test keeps its per-atom energies
    if (j.kind == POT_NNP && (rc = plug_put(h, L.en_atom, new_en_atom, C * N * 8))) return rc;
    j.x = A + L.x;
    j.d2_in = d2 ? A + L.d2 : nullptr;
    j.box = box ? A + L.box : nullptr;
    j.index = reinterpret_cast<const int32_t *>(A + L.index);
    j.trial = A + L.trial;
    j.en_in = A + L.en_in;
    j.vars_in = nv ? A + L.vars_in : nullptr;
    j.en = A + L.en;
    j.en_atom = A + L.en_atom;
    j.vars_out = nv ? A + L.vars_out : nullptr;
    j.ws = A + L.ws;
    HIPCHK(h, plugin_update(j, h->stream));
    if ((rc = plug_get(h, new_en, L.en, C * 8)) ||
        (rc = plug_get(h, nv ? new_en_atom : nullptr, L.en_atom, C * N * 8)) ||
        (rc = plug_get(h, nv ? new_vars : nullptr, L.vars_out, C * (size_t)nv * 8)))
        return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return PTMC_OK;
}

}  // extern "C"
