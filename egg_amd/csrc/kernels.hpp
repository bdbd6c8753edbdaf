// Device-side data descriptors and kernel launchers of the structured TMOP
// smoother. Host code (session.cpp) fills these and calls the launchers; the
// kernels live in sweep_kernels.hip.
//
// Store layout (HBM): the halo-padded per-block node slots of the reference
// (src/structured.hpp:47-163 BlockLayout, row-major padded axes, blocks back
// to back, overflow slots last) with the coordinate axis split into planes:
// coordinate k of slot n is x[k * plane + n]. Node slot numbers are therefore
// identical to the reference's real-offset / D, so every index table (halo,
// share, fan, patch stencils, g2s) is the reference's table in node units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace eggmi
{

struct BlockDesc {
    long long base;  // node slot of padded index (0, ..., 0)
    int n[3];        // interior node counts per axis
    int pad;
    long long st[3];  // node strides of the padded axes (last axis = 1)
};

// A tile carries its block's geometry (one 64-byte load instead of two
// dependent ones at the head of the latency-bound small-grid launches).
struct __align__(16) TileDesc {
    BlockDesc b;
    int i0, j0, k0;  // first logical index of the tile on each axis
    int blk;
};

// Stored-stencil DOFs (block faces, interfaces, singular corners). A group
// of G lanes serves one DOF (G = stored_lanes(dim) by default; some tile
// variants take fewer lanes and more rounds), one occurrence per lane and
// round: DOF d's occurrence r * G + l sits at padded position
// (d * rounds + r) * G + l, so a lane finds its occurrence without reading
// an offset table. Padding entries carry meta bit 7.
__host__ __device__ constexpr int stored_lanes(int dim) { return dim == 2 ? 16 : 64; }
constexpr uint8_t kStoredPad = 0x80;
struct StoredView {
    int n = 0;
    int rounds = 1;             // ceil(max occurrences of a DOF / lanes)
    const int* dof = nullptr;   // [n] slot written
    long long total = 0;        // padded occurrences: n * rounds * lanes
    const int* gc = nullptr;    // [total] corner slot
    const int* gn = nullptr;    // [dim][total] neighbour slots
    const uint8_t* meta = nullptr;  // [total] bits 0-2: role+1 (0..D), bits 3..5: sign of axis k (1 = -1), bit 7: padding
    const float* W = nullptr;       // [total][dim*dim] per-occurrence W_inv rows, or null (uniform)
    const float* wt = nullptr;      // [total] per-occurrence energy weight, or null (1)
};

constexpr int kCounterSlots = 64, kCounterStride = 16;

// Directional soft-energy samples on the store: node slots (4 per sample,
// the first arity of them used), kind (0 parallel, 1 line, 2 stem, -1
// none), frozen reference vector (D per sample), weight. As a per-DOF
// window table, n is the windows per DOF and slot the DOF's place (-1 pads);
// as the energy reports' whole-sample table, n is the sample count.
struct DirView {
    int n = 0;
    float eps = 0.0f;
    const int* nodes = nullptr;
    const int8_t* slot = nullptr;
    const int8_t* kind = nullptr;
    const float* ref = nullptr;
    const float* w = nullptr;
};

struct SweepArgs {
    const float* xin = nullptr;
    float* xout = nullptr;
    long long plane = 0;
    const TileDesc* tiles = nullptr;
    int n_tiles = 0;
    const BlockDesc* blocks = nullptr;
    const uint8_t* mask = nullptr;  // [slots] 1 = synthesized interior DOF
    StoredView st;
    float omega = 1.0f;
    float delta = 0.0f;
    const float* tau = nullptr;  // FAS linear shift (plane layout), null = none
    // constrained boundary DOFs: stored DOFs st.n .. st.n + con_n - 1 of the
    // stored arrays, each sliding on the entity con_tag[i] whose blob
    // (entity.hpp) starts at con_blob + con_off[i]; con_seed[2 i ..] the
    // foot parameter of its last projection (NaN: none yet), con_epoch[i]
    // the run serial of its last update (*run_epoch: the current run's)
    int con_n = 0;
    const uint8_t* con_tag = nullptr;  // 0: a free DOF (there for its curvature windows)
    const float* con_blob = nullptr;
    const int* con_off = nullptr;
    float* con_seed = nullptr;
    int* con_epoch = nullptr;
    const int* run_epoch = nullptr;
    // block-interface C2 curvature windows of those DOFs (2D): curv_k per DOF,
    // four node slots each, the DOF's place (-1 = padding) and the weight
    int curv_k = 0;
    const int* curv_nodes = nullptr;
    const int8_t* curv_slot = nullptr;
    const float* curv_w = nullptr;
    // directional soft-energy windows of those DOFs: dir.n per DOF
    DirView dir;
    float* gout = nullptr;       // gradient-only launches: per-node gradient (plane layout)
    bool gsub = false;           // gradient-only: gout = gradient - gout, in place
    bool gpush = false;          // gradient-only: also write it into the DOF's replica slots
    const int* cp_off = nullptr;  // [slots+1] replicas of each slot (push copies), null = none
    const int* cp_dst = nullptr;
    float Wu[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // uniform W_inv for stored occurrences
    // optional counters (null = off): kCounterSlots rows of kCounterStride
    // words, [0] line-search trials evaluated, [1] DOF updates, [2] accepted
    // learned proposals; the host sums the rows
    unsigned long long* counters = nullptr;
    // learned proposal for synthesized interior DOFs (fwd == null: off);
    // mode 0: an accepted proposal replaces the Newton step, 1: the Newton
    // step continues from it; the network's displacement is scaled by
    // prop_scale
    MlpView mlp;
    int prop_mode = 0;
    float prop_scale = 1.0f;
};

// ------------------------------------------------------------------ multi-sweep
// k block-Jacobi sweeps in one launch (2D levels without special DOFs). A
// workgroup owns a core tile of DOFs and stages in LDS its dependency
// closure: every node the core's k-th sweep depends on, k rings deep through
// block interfaces, replicas and stored stencils, as rectangles of block
// padded boxes (regions: synthesized DOFs use the 3x3 row addressing of the
// one-sweep kernel) plus single extra nodes. Sub-sweep s recomputes every
// DOF within distance k-1-s of the core from the LDS state of sub-sweep
// s-1 (the outer rings redundantly, as the neighbouring tiles do for theirs),
// ping-ponging two LDS buffers; the last sub-sweep writes the core to the
// store and pushes its replicas. Per DOF the arithmetic is the one-sweep
// kernel's device code (the compiler may fuse its multiply-adds differently
// in the two kernels: results agree to a few float ulps).
constexpr int kMsMaxK = 8;
struct __align__(16) MsRegion {
    long long base;  // store slot of the rectangle's first node
    long long st0;   // slot stride between its rows (the block's padded row)
    int h, w;        // rows, columns
    int lds;         // LDS entry of the first node (row-major)
    int pad;
};
struct __align__(16) MsTile {
    int reg0, nreg;  // regions
    int ext0, next;  // extra nodes (store slots), LDS entries n_lds - next .. n_lds - 1
    int n_lds;
    int syn0;               // first synthesized entry of the tile (entries sorted by distance)
    int syn_cnt[kMsMaxK];   // synthesized entries within distance r of the core
    int st0;                // first stored entry of the tile (sorted by distance)
    int st_cnt[kMsMaxK];    // stored entries within distance r of the core
    int push0, npush;       // LDS replica pairs (a ghost in a region <- its root)
};
struct MsArgs {
    int k = 0;           // sweeps per launch (<= the depth the tables were built for)
    int n_tiles = 0;
    int lds_stride = 0;  // LDS floats per coordinate plane of one buffer (>= every tile's n_lds)
    const MsTile* tiles = nullptr;
    const MsRegion* regions = nullptr;
    const int* ext = nullptr;       // extra nodes' store slots
    // synthesized entries: x = LDS entry | (its region's row width << 16), y = store slot
    const int2* syn = nullptr;
    const int* st_slot = nullptr;   // store slot of each stored entry
    StoredView st;                  // stored entries: dof, gc, gn are LDS entries
    const int* push_src = nullptr;  // LDS entries
    const int* push_dst = nullptr;
};
// dynamic LDS of a multi-sweep launch: two buffers of D planes
inline size_t ms_lds_bytes(int dim, int stride) { return (size_t)stride * 2 * dim * sizeof(float); }
constexpr size_t kMsMaxLdsBytes = 64 * 1024;
// kernel variants: lanes per node (4 / CPL) and workgroup size
struct MsVariant {
    int cpl, threads;
};
int num_ms_variants();
MsVariant ms_variant(int v);
// k sweeps (m.k) in one launch; a.xin holds the state on entry, a.xout gets
// the core DOFs (and their replicas) k sweeps later
void launch_multi_sweep(int kind, int variant, const SweepArgs& a, const MsArgs& m, hipStream_t s);

// Energy / min-det reduction over the structured cell stencil.
struct CellReduceArgs {
    const float* x = nullptr;
    long long plane = 0;
    const BlockDesc* blocks = nullptr;
    const long long* cell_off = nullptr;  // [nblocks+1] exclusive prefix of cells per block
    int nblocks = 0;
    long long ncells = 0;
    double delta = 0.0;
    const float* tau = nullptr;  // optional: subtract sum tau.x over owner slots (FAS)
};

// Energy / min-det reduction over an explicit sample table (fallback when the
// wire's energy stencil is not the plain structured cell stencil).
struct SampleReduceArgs {
    const float* x = nullptr;
    long long plane = 0;
    long long ns = 0;
    const int* gc = nullptr;
    const int* gn = nullptr;  // [dim][ns]
    const float* s = nullptr;  // [dim][ns]
    const float* W = nullptr;  // [ns][dd] or null (uniform)
    const float* wt = nullptr;  // [ns] energy weight, or null (1)
    double Wu[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double delta = 0.0;
};

constexpr int kReduceBlocks = 1024;
constexpr int kReduceThreads = 256;

// Launchers (stream-ordered, asynchronous). kind: 0 shape, 1 untangle, 2 shape_size.
void launch_sweep(int dim, int kind, int tile_variant, const SweepArgs& a, hipStream_t s);
// Metric pass only: per-DOF gradient (minus tau when given) into a.gout
// (a.gsub: a.gout = gradient - a.gout, in place).
void launch_grad(int dim, int kind, int tile_variant, const SweepArgs& a, hipStream_t s);
// n_sweeps tau-shifted sweeps of a FAS coarse level (synthesized DOFs only)
// in one single-workgroup launch, over the variant's tile table: x0 holds
// the state on entry (a.xin / a.xout are replaced), the result ends in x0
// for an even n_sweeps and in x1 for an odd one.
void launch_level_smooth(int dim, int kind, int tile_variant, const SweepArgs& a, float* x0, float* x1,
                         int n_sweeps, hipStream_t s);
void launch_copy(float* x, long long plane, int dim, const int* src, const int* dst, long long n, hipStream_t s);
// *p = v, stream-ordered (a kernel, so it may sit in captured work)
void launch_set_int(int* p, int v, hipStream_t s);
void launch_reduce_cells(int dim, int kind, const CellReduceArgs& a, double* partials, hipStream_t s);
void launch_reduce_samples(int dim, int kind, const SampleReduceArgs& a, double* partials, hipStream_t s);
// Folds the partials into *out_e, *out_m; with rep set, also appends them
// as report *count (rep[*count] and rep[cap + *count], when *count < cap)
// and increments *count.
void launch_reduce_finish(const double* partials, int nparts, double* out_e, double* out_m, hipStream_t s,
                          double* rep = nullptr, int cap = 0, int* count = nullptr);
// The whole-sample directional energies in FP64 at x + alpha_t corr for
// mode 0: alpha 0 (the energy reports; corr may be null), 1: a0, a1 (the
// line-search pair), 2: the FAS ladder's slots; written as the partials of
// blocks pbase, pbase + 1, ... in the layout of that mode's reduction (2,
// 4 or 2 (kLadder + 1) doubles per block, min det +inf). Returns the block
// count (at most kReduceBlocks / 2), 0 when the table is empty.
int launch_dir_reduce(int dim, int mode, const DirView& v, const float* x, long long plane, const float* corr,
                      float a0, float a1, double* partials, int pbase, hipStream_t s);

// Tile shapes the sweep kernels are instantiated for (the autotuner picks one).
struct TileShape {
    int tx, ty, tz;
    int sg = 0;  // lanes per stored DOF (0 = stored_lanes(dim))
};
int num_tile_variants(int dim);
TileShape tile_variant_shape(int dim, int v);
int tile_variant_stored_lanes(int dim, int v);
// whether the variant has a learned-proposal kernel (the 2D one-lane-per-
// sample variants do not)
bool tile_variant_has_proposal(int dim, int v);
// the variant whose proposal kernel runs on variant v's tile table (same
// tile shape and stored lanes): v itself when it has one, else -1 if none
int tile_variant_proposal_partner(int dim, int v);

}  // namespace eggmi
