// Host side of the MI355X structured smoother: wire validation, the
// structured remap onto the halo-padded store, device residency and the
// sweep / FAS drivers. Plain C++ over the HIP runtime; exposed through the C
// ABI in capi.cpp (include/eggmi.h).
#pragma once

#include "control.hpp"
#include "eggmi.h"
#include "kernels.hpp"
#include "mg.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace eggmi
{

struct InvalidArgument : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Internal : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);
#define EGGMI_HIP(expr) ::eggmi::hip_check((expr), #expr)

// Owning device array.
template <class T> struct DArr {
    T* p = nullptr;
    size_t n = 0;
    DArr() = default;
    DArr(const DArr&) = delete;
    DArr& operator=(const DArr&) = delete;
    DArr(DArr&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DArr& operator=(DArr&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DArr() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        release();
        n = count;
        if (count) EGGMI_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(const std::vector<T>& h)
    {
        alloc(h.size());
        if (n) EGGMI_HIP(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
    }
};

// Per-block geometry of one level of the halo-padded store
// (reference src/structured.hpp:47-163 BlockLayout<D>, in node units).
struct Layout {
    int dim = 2;
    std::vector<std::array<int, 3>> shape;  // interior node counts
    std::vector<long long> base;            // node slot of padded (0,..,0)
    std::vector<std::array<long long, 3>> stride;
    long long padded_nodes = 0;  // sum of padded block sizes
    long long overflow = 0;      // extra slots past the last block

    void build(int d, const std::vector<std::array<int, 3>>& shapes);
    long long n_slots() const { return padded_nodes + overflow; }
    long long padded_slot(int b, const int* padded) const;
    long long interior_slot(int b, const int* logical) const;
    long long block_padded_count(int b) const;
};

// Cross-block copy tables of one level in block / padded-index form (face
// ghost fills) plus owner -> copy broadcasts (slots): the input of the coarse
// interface table derivation.
struct HaloHost {
    std::vector<int> sb, db;
    std::vector<std::array<int, 3>> sp, dp;
    std::vector<long long> share_src, share_dst;
};

struct FasParams {
    int nu_pre = 2, nu_post = 2, nu_coarse = 32;
    float omega = 1.0f;
};

// Everything one level of the hierarchy needs on the device.
struct Level {
    Layout lay;
    long long plane = 0;  // coordinate plane stride (slots rounded up)
    DArr<BlockDesc> blocks;
    // fused copy table: halo [0, n_halo), fan [n_halo, n_halo+n_fan), share after
    std::vector<int> copy_src, copy_dst;
    long long n_halo = 0, n_fan = 0, n_share = 0;
    DArr<int> d_copy_src, d_copy_dst;
    // push form of the copies whose (root) source is a relaxed DOF: the
    // sweep kernel writes a DOF's new value into every replica slot too
    std::vector<int> cp_off, cp_dst;
    DArr<int> d_cp_off, d_cp_dst;
    long long n_push = 0;
    // every copy in push form (any root source, relaxed or not) and a mask
    // of the copy destinations: a transfer kernel writes a node's value into
    // its replicas itself, with no halo launch after it
    std::vector<int> rp_off, rp_dst;
    std::vector<uint8_t> is_dst;
    DArr<int> d_rp_off, d_rp_dst;
    DArr<uint8_t> d_is_dst;
    // synthesized interior DOFs
    std::vector<uint8_t> mask;  // [n_slots]
    DArr<uint8_t> d_mask;
    long long n_synth = 0;
    std::vector<TileDesc> tiles;
    DArr<TileDesc> d_tiles;
    int variant = 0;  // the tile variant the tile table is laid out for
    bool one_wg = false;  // (FAS coarse level) each smoothing call is one single-workgroup launch
    std::vector<int> syn_slot, syn_blk;  // synthesized DOFs as a list (lazily, for harvesting)
    DArr<int> d_syn_slot, d_syn_blk;
    void ensure_synth_list();
    // stored-stencil DOFs
    std::vector<int> st_dof, st_off, st_gc, st_gn;
    std::vector<uint8_t> st_meta;
    std::vector<float> st_W;  // empty = uniform
    std::vector<float> st_wt;  // per-occurrence energy weight; empty = 1
    // the last con_tag.size() stored DOFs are special: constrained to the
    // entity con_tag (0: none, the DOF is there for its windows) whose FP32
    // record (entity.hpp) starts at con_blob[con_off]; per DOF the foot
    // parameter of its last projection (NaN: none, the iterative
    // projections' warm start) and the run serial of its last update
    // against the level's current one (a DOF's first update in a run is
    // cold)
    std::vector<uint8_t> con_tag;
    std::vector<int> con_off;
    std::vector<float> con_blob;
    DArr<uint8_t> d_con_tag;
    DArr<int> d_con_off;
    DArr<float> d_con_blob;
    DArr<float> d_con_seed;
    DArr<int> d_con_epoch;
    DArr<int> d_run_epoch;
    int run_epoch = 0;
    void reset_seeds();              // positions replaced: no warm start
    void next_run(hipStream_t s);    // a new run begins on stream s
    // the special DOFs' seeds and run serials, kept across timing runs
    struct ConState {
        std::vector<float> seed;
        std::vector<int> epoch;
        int run_epoch = 0;
    };
    ConState save_con_state() const;
    void restore_con_state(const ConState& c);
    // their curvature windows (curv_k per special DOF): owner slots of the
    // four nodes, the DOF's place in the window (-1 = padding), the weight
    int curv_k = 0;
    std::vector<int> curv_nodes;
    std::vector<int8_t> curv_slot;
    std::vector<float> curv_w;
    DArr<int> d_curv_nodes;
    DArr<int8_t> d_curv_slot;
    DArr<float> d_curv_w;
    // their directional windows (dir_k per special DOF): node slots (4 per
    // window), the DOF's place (-1 = padding), kind (-1 = none), reference
    // vector (2), weight; one eps
    int dir_k = 0;
    float dir_eps = 0.0f;
    std::vector<int> dir_nodes;
    std::vector<int8_t> dir_slot, dir_kind;
    std::vector<float> dir_ref, dir_w;
    DArr<int> d_dir_nodes;
    DArr<int8_t> d_dir_slot, d_dir_kind;
    DArr<float> d_dir_ref, d_dir_w;
    // the energy reports' whole-sample directional table (same layout, no
    // slot) and its device view (n = 0: none)
    std::vector<int> es_dir_nodes;
    std::vector<int8_t> es_dir_kind;
    std::vector<float> es_dir_ref, es_dir_w;
    DArr<int> d_es_dir_nodes;
    DArr<int8_t> d_es_dir_kind;
    DArr<float> d_es_dir_ref, d_es_dir_w;
    DirView es_dir;
    float Wu[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    DArr<int> d_st_dof, d_st_gc, d_st_gn;  // device copies in the padded lane layout (see StoredView)
    int st_rounds = 1;
    long long st_padded = 0;
    int st_lanes = 0;  // lanes per stored DOF of the uploaded layout (0 = none yet)
    DArr<uint8_t> d_st_meta;
    DArr<float> d_st_W, d_st_wt;
    // energy reduction: structured cells, or an explicit sample table
    bool structured_energy = true;
    std::vector<long long> cell_off;
    DArr<long long> d_cell_off;
    long long ncells = 0;
    long long es_n = 0;
    DArr<int> d_es_gc, d_es_gn;
    DArr<float> d_es_s, d_es_W, d_es_wt;  // d_es_wt: per-sample energy weight, empty = 1
    double es_Wu[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    // device state
    DArr<float> x[2];
    int cur = 0;  // which x holds the canonical state
    // FAS state (coarse levels; the fine level uses grad / corr / d_free)
    std::vector<std::array<int, 3>> factor;  // per block, vs the parent level
    std::vector<uint8_t> free_mask, free_interior;
    DArr<uint8_t> d_free;
    DArr<int> d_factor;
    DArr<long long> d_coarse_off, d_fine_off;  // node-count prefixes for the transfers
    long long n_coarse_nodes = 0, n_fine_nodes = 0;
    DArr<float> rx, tau, grad, corr;  // corr: the fine level's correction only

    // multi-sweep tables (multisweep.cpp; 2D): k sweeps per launch over
    // LDS-staged dependency closures of core tiles of ty x tx logical nodes
    struct MsHost {
        int k = 0, ty = 0, tx = 0, lds_stride = 0, rounds = 1, G = 16;
        std::vector<MsTile> tiles;
        std::vector<MsRegion> regions;
        std::vector<int> ext, st_slot, st_lds, gc, gn, push_src, push_dst;
        std::vector<int2> syn;  // synthesized entries (MsArgs::syn)
        std::vector<uint8_t> meta;
        std::vector<float> W;  // per-occurrence W_inv rows (empty: uniform)
        long long n_st = 0, total = 0;
        long long core_dofs = 0, updates = 0;  // DOFs, and DOF updates per launch over its k sub-sweeps
    };
    MsHost ms;  // ms.k == 0: none built
    int ms_kernel = 0;  // kernel variant (ms_variant)
    int ms_sweeps = 0;  // sweeps per launch in use (0: the one-sweep kernels)
    DArr<MsTile> d_ms_tiles;
    DArr<MsRegion> d_ms_regions;
    DArr<int2> d_ms_syn;
    DArr<int> d_ms_ext, d_ms_st_slot, d_ms_st_lds, d_ms_gc, d_ms_gn, d_ms_push_src, d_ms_push_dst;
    DArr<uint8_t> d_ms_meta;
    DArr<float> d_ms_W;
    // builds the tables for depth k (throws InvalidArgument when the level
    // is not eligible or a closure outgrows the LDS) and uploads them
    void build_multi(int k, int ty, int tx, bool upload);
    MsArgs multi_args(int k) const;
    // max |difference| between k sweeps of a synthetic nonlinear update run
    // through the tables' schedule and k plain synchronous sweeps of it
    double multi_selfcheck(int k, unsigned seed) const;

    BlockDesc block_desc(size_t b) const;
    void upload_tables();
    void make_tiles(int v);  // also lays the stored DOFs out for the variant's lane groups
    void upload_stored(int G);
    void build_push();
    SweepArgs sweep_args(const float* xin, float* xout, float omega, float delta, const float* tau) const;
};

struct Mlp;

// Control-net reduced Gauss-Newton state of one block (control_host.cpp).
struct ControlOptions {
    int n_outer = 30;
    double grad_tol = 1e-8, alpha_min = 1e-4, lm0 = 0.0;
    int lm_tries = 8, pcg_max_iter = 200;
    double pcg_rtol = 1e-10;
    bool pcg_forcing = true;
};
struct ControlReport {
    std::vector<double> energies, mindets, alphas, pcg_iters, pcg_relres;
    int iters = 0;
    bool converged = false;
};
struct ControlNet {
    CtrlGeom g;  // sizes, block geometry and the device tables below
    int block = 0;
    DArr<int> off[3], sup_off[3], sup_node[3], free_idx, l2f;
    DArr<double> wt[3], sup_w[3], C, b;
    // per sample: metric gradient and symmetric Hessian; block-local node
    // fields (positions, step, gathers); free-control vectors; scalars
    DArr<double> gs, zs, hs, xb, ynode, dir, hnode, prec, G, dC, r, z, p, Ap, scal, eparts;
    // the other blocks' cells (their energy and min det, constant during a run)
    DArr<BlockDesc> oblocks;
    DArr<long long> ocell_off;
    int n_oblocks = 0;
    long long n_ocells = 0;
};

// A net's per-axis windows (checked to stay inside the lattice, finite)
// and their inverse (control -> nodes with weights) onto the device; g
// holds the node and lattice shapes and the window width.
void upload_net_axes(int dim, const CtrlGeom& g, const int32_t* const* off, const double* const* wt, DArr<int>* d_off,
                     DArr<double>* d_wt, DArr<int>* d_soff, DArr<int>* d_snode, DArr<double>* d_sw);

// Several blocks' nets over one stacked control vector with a reduction
// dC = R dq to the reduced DOFs and homogeneous penalty rows
// (control_topo.cpp; reference ControlTopoSessionT,
// src/control_solve.hpp:1585-2837).
struct ControlTopo {
    int n_nets = 0;
    std::vector<int> net_block;
    std::vector<long long> net_nodes;  // node count of each net's block
    std::vector<TopoBlock> hblk;
    std::vector<std::array<DArr<int>, 3>> off, sup_off, sup_node;
    std::vector<std::array<DArr<double>, 3>> wt, sup_w;
    std::vector<DArr<double>> b;
    DArr<TopoBlock> blk;
    DArr<long long> npfx, spfx, cpfx;
    DArr<int> net_blk, owner, grp_off, grp_mem;
    TopoView view;
    long long n_all = 0, ns_all = 0, n_stack = 0, nq = 0, n_pen = 0;
    // R (n_stack * dim rows, nq columns) and its transpose (rows ascending
    // within a column); penalty rows P over the stacked components, their
    // weights and transpose
    DArr<int> r_off, r_col, g_off, g_row, p_off, p_col, pt_off, pt_row;
    DArr<double> r_coef, g_coef, p_coef, p_w, pt_coef, r_pen;
    std::vector<double> p_w_h;
    // stacked lattice; node fields over every block node; per-sample
    // derivatives; stacked vectors; reduced-space PCG vectors; scalars
    DArr<double> C, xb, dir, ynode, hnode, gs, hs, zs, prec, vfull, wfull, dCfull;
    DArr<double> qG, qdq, qr, qz, qp, qAp, qdiag, scal, eparts;
};

struct Session {
    int dim = 2;
    long long M = 0;  // global node count
    std::vector<int> g2s;
    std::vector<std::vector<int>> bgd;  // per-block global ids (row-major)
    Level fine;
    int tile_variant = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DArr<double> partials;
    DArr<double> d_reports;
    double last_ms = 0.0;
    long long last_launches = 0;
    long long last_fas_graph_kernels = 0;  // kernel nodes of the last replayed FAS cycle graph
    bool kernel_timing = false;     // bracket every sweep launch with events
    // hipGraph replay of sweep chunks (graph_chunk sweeps, even; 0 = off):
    // one captured graph per starting buffer, re-captured when a parameter
    // baked into it changes
    struct SweepGraph {
        hipGraphExec_t exec = nullptr;
        int kind = -1, variant = -1, mode = 0, chunk = 0, every = 1;
        float omega = 0.0f, delta = 0.0f, scale = 0.0f;
        const void* mlp = nullptr;
    };
    int graph_chunk = 0;
    SweepGraph graphs[2];
    void release_graphs();
    hipGraphExec_t sweep_graph(int parity, int kind, float omega, float delta, bool propose);
    double last_kernel_ms = 0.0;    // summed sweep-kernel time of the last run
    std::vector<hipEvent_t> tev;
    void ensure_timing_events(size_t n);
    // diagnostic counters per level: [4*level] trials evaluated, [4*level+1]
    // DOF updates, [4*level+2] accepted learned proposals
    static constexpr int kMaxCounterLevels = 32;
    static constexpr int kCountersPerLevel = 4;
    bool counting = false;
    DArr<unsigned long long> d_counters;
    std::vector<long long> last_counters;
    unsigned long long* counters_for(const Level& L);
    void counters_begin();
    void counters_end();
    // FAS hierarchy (built lazily)
    std::vector<std::unique_ptr<Level>> coarse;
    int built_levels = -1;
    // tile variant of each coarse level (-1 = the session's tile_variant)
    std::vector<int> level_variant;
    // level variant of coarse level i: a tile variant v < V launches every
    // sweep over the level's tiles; V + v smooths the whole level in one
    // single-workgroup launch per smoothing call on variant v's tiles
    // (V = num_tile_variants)
    int coarse_variant(size_t i) const;
    void apply_level_variant(size_t i);
    std::vector<uint8_t> fine_free;  // relaxed fine DOF slots
    HaloHost fine_halo;
    DArr<double> d_ls;
    bool host_only = false;  // created with device = 3: index tables only
    bool weighted = false;   // the wire carries per-sample energy weights other than 1
    const Mlp* proposal = nullptr;  // learned proposal for run_jacobi (not owned)
    int proposal_mode = 0;
    float proposal_scale = 1.0f;
    // cadence: plain runs propose in sweeps 0, k, 2k, ... of each run call;
    // with k > 1 a FAS fine smoothing call proposes in its first sweep only
    int proposal_every = 1;
    bool proposes_at(int sweep) const { return sweep % proposal_every == 0; }
    // the variant the proposal sweeps launch (same tile table as tile_variant)
    int proposal_variant() const { return tile_variant_proposal_partner(dim, tile_variant); }
    bool pair_valid = false;        // the second Jacobi buffer holds the previous sweep's state
    // control-net reduced Gauss-Newton (control_host.cpp)
    std::unique_ptr<ControlNet> control;
    void set_control(const eggmi_control& c);
    void control_set_C(const double* C, long long n);
    void control_get_C(double* C, long long n) const;
    void control_eval_to_x();
    ControlReport run_control(const ControlOptions& o, int kind);
    // several blocks' nets (control_topo.cpp); set_C / get_C / run_control
    // act on the stacked vector while one is attached
    std::unique_ptr<ControlTopo> control_topo;
    void set_control_topo(const eggmi_control_topo& t);
    void control_topo_set_reduction(long long nq, const int* off, const int* col, const double* coef);
    void control_topo_set_penalty(long long n_pen, const int* off, const int* col, const double* coef,
                                  const double* w);
    void control_topo_set_b(int net, const double* b, long long n);
    void control_topo_eval_to_x();
    ControlReport run_control_topo(const ControlOptions& o, int kind);

    Session() = default;
    ~Session();

    void create(int dim, const eggmi_group* groups, int n_groups, const eggmi_stencil* es,
                const eggmi_structured* st, const double* X, long long num_nodes, int device);
    void halo(Level& L, float* x);
    // energy and min det of x into out_e / out_m (device); with rep set,
    // also appended as report *count of the capacity-cap report buffer rep
    void reduce(Level& L, const float* x, int kind, double delta, double* out_e, double* out_m,
                double* rep = nullptr, int cap = 0, int* count = nullptr);
    std::pair<std::vector<double>, std::vector<double>> run_jacobi(int n_sweeps, int kind, double delta,
                                                                   double omega, int report_every);
    std::pair<std::vector<double>, std::vector<double>> run_fas(int n_cycles, int nu_pre, int nu_post,
                                                                int nu_coarse, double omega, int max_levels,
                                                                int kind);
    std::vector<std::vector<std::array<int, 3>>> mg_level_shapes(int max_levels);
    void get_X(double* out) const;
    void set_X(const double* X);
    std::vector<float> packed() const;
    void need_device() const;
    void set_tile_variant(int v);
    // FP64 energy and min det of the resident state (one reduction)
    std::pair<double, double> energy(int kind, double delta);
    // time n_sweeps plain sweeps per tile variant from the current state
    // (restored after each trial), keep the fastest; returns ms per variant
    std::vector<double> autotune(int n_sweeps, double omega);
    // time n_sweeps plain sweeps per multi-sweep candidate (k, ty, tx,
    // kernel) on the fine level from the current state (restored after each
    // trial); keeps the fastest if it beats base_ms (else the previous
    // setting); returns ms per candidate (+inf: not buildable)
    std::vector<double> autotune_multi(int n_sweeps, double omega, const std::vector<std::array<int, 4>>& cands,
                                       double base_ms);
    // per FAS coarse level (from the fine state restricted down the chain):
    // time one smoothing call of n_sweeps as the level's one-sweep setting
    // and per multi-sweep candidate, keep the fastest; ms [level][1 + cand]
    std::vector<double> autotune_levels_multi(int n_sweeps, double omega, int max_levels, int kind,
                                              const std::vector<std::array<int, 4>>& cands);
    // k sweeps per launch on level `level` (0 = fine, i >= 1 = FAS coarse
    // level i, which prepare_fas must have built) over ty x tx core tiles
    // with multi-sweep kernel variant `kernel`; k <= 1 returns the level to
    // the one-sweep kernels (multisweep.cpp)
    void set_multi_sweep(int level, int k, int ty, int tx, int kernel);
    Level& level_ref(int level);
    // k of the launches sweeping L (0 = one-sweep kernels); sweeps with the
    // learned proposal keep the one-sweep kernels
    static int multi_k(const Level& L, bool propose) { return L.ms_sweeps > 1 && !propose ? L.ms_sweeps : 0; }
    // launches in a captured chunk of graph_chunk sweeps
    int chunk_launches(int mk) const { return mk ? (graph_chunk + mk - 1) / mk : graph_chunk; }
    // one multi-sweep launch of k sweeps on L: xin -> xout
    void launch_multi(Level& L, const float* xin, float* xout, int k, int kind, float omega, float delta,
                      const float* tau);
    // per-level tile variants of the FAS coarse levels (-1 = tile_variant)
    void set_level_variants(const std::vector<int>& v);
    // time n_sweeps FAS-shifted sweeps per tile variant on each of the first
    // max_levels - 1 coarse levels, from the fine state restricted down the
    // chain, and keep each level's fastest; returns ms [level][variant]
    std::vector<double> autotune_levels(int n_sweeps, double omega, int max_levels, int kind);
    // FAS internals
    void build_hierarchy();
    // transfer operators of a level pair applied to caller fields (fas.cpp)
    void mg_probe(int level, int op, const double* in, double* out);
    size_t prepare_fas(int max_levels);  // device state of the levels a cycle uses; their count
    void upload_level(size_t i);
    TransferArgs transfer(const Level& parent, const Level& child, bool prolong) const;
    void smooth(Level& L, int sweeps, int kind, float omega, const float* tau);
    // per-DOF (tau-shifted) gradient into out; sub: out = gradient - out;
    // push: also into the DOFs' replica slots
    void grad(Level& L, const float* x, float* out, const float* tau, int kind, bool sub = false,
              bool push = false);
    void restrict_to(Level& parent, const float* x_parent, Level& child, int kind);
    void coarse_vcycle(size_t li, size_t n_levels, int kind, const FasParams& o);
    // one FAS cycle in three parts: pre-smoothing to the halo-consistent
    // fine correction; the device line-search ladder and application; post-
    // smoothing and the report reduction
    void fas_head(int kind, const FasParams& o, size_t n_coarse);
    void fas_device_ladder(int kind);
    void fas_tail(int kind, const FasParams& o, double* out_e, double* out_m, double* rep = nullptr, int cap = 0,
                  int* count = nullptr);
    CellReduceArgs fine_cells(const float* x) const;
    // a whole cycle captured as a hipGraph, per parity of the fine Jacobi
    // buffer it starts from (the ladder decided on the device, the report
    // appended to d_reports at the device counter d_rep_count)
    struct FasGraph {
        hipGraphExec_t exec = nullptr;
        int kind = -1, variant = -1, nu_pre = 0, nu_post = 0, nu_coarse = 0, mode = 0;
        size_t n_coarse = 0;
        float omega = 0.0f, scale = 0.0f;
        const void* mlp = nullptr;
        const double* rep = nullptr;
        int every = 1;
        long long launches = 0;
        long long kernels = 0;  // kernel nodes
    };
    FasGraph fas_graphs[2];
    hipGraphExec_t fas_graph(int parity, int kind, const FasParams& o, size_t n_coarse);
    DArr<double> d_fas_reports, d_ladder_parts, d_ladder_out, d_cyc_rep;
    DArr<float> d_alpha;
    DArr<int> d_rep_count;
};

}  // namespace eggmi
