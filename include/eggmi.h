/*
 * eggmi — C ABI of the MI355X structured TMOP smoother / untangler.
 *
 * Drop-in compute core for the reference's pybind module egg._cpp.cpp_core
 * (reference src/bindings.cpp:1541-2130, stub egg/_cpp/cpp_core.pyi:27-68).
 * Every pointer is borrowed for the duration of the call and owned by the
 * caller; the library copies what it keeps. Functions return an eggmi_status;
 * on failure eggmi_last_error() returns the message of the calling thread's
 * last error. A session is not thread-safe (the reference holds the GIL for
 * every call, src/bindings.cpp has no gil_scoped_release).
 */
#ifndef EGGMI_H
#define EGGMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EGGMI_OK = 0,
    EGGMI_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument -> ValueError */
    EGGMI_NO_DEVICE = 2,        /* reference: select_queue runtime_error -> RuntimeError */
    EGGMI_INTERNAL = 3          /* reference: std::logic_error / device failure */
} eggmi_status;

typedef enum { EGGMI_PHASE_BARRIER = 0, EGGMI_PHASE_UNTANGLE = 1, EGGMI_PHASE_SHAPE_SIZE = 2 } eggmi_phase;

typedef struct eggmi_session eggmi_session;

/* One sweep group of the flat wire (reference egg/smoothing/flat_context.py:
 * 283-300, unpacked at src/bindings.cpp:334-490). Global node ids. */
typedef struct {
    int64_t ndof;                     /* "D" */
    const int32_t* dof_idx;           /* [ndof] */
    const int32_t* P_of;              /* [ndof] stored occurrences per DOF */
    const int32_t* interior_block;    /* [ndof] block of a synthesized interior DOF, -1 = stored */
    const int32_t* interior_logical;  /* [ndof*dim] */
    int64_t total_samples;            /* sum(P_of) */
    const int32_t* role;              /* [total] -1 absent, 0 corner, 1..dim neighbour */
    const int32_t* gc;                /* [total] */
    const int32_t* gn[3];             /* [total] each, gn[0..dim-1] */
    const double* s[3];               /* [total] each, +-1 */
    const double* W_inv;              /* [w_rows*dim*dim] */
    int64_t w_rows;                   /* 1 (uniform target) or total */
    const double* weight;             /* [total] per-occurrence energy weight (scales mu's value,
                                         gradient and Hessian, not det), NULL = 1
                                         (src/patch.hpp:45-77, :250-292) */
    const int32_t* entity_tag;        /* [ndof] entity kind a boundary DOF slides on, the
                                         reference's tags (src/entity_soa.hpp:37-55): 0 free,
                                         2D 1 segment, 2 circle, 3 ellipse, 6 circle arc,
                                         7 ellipse arc, 8 / 9 quadratic / cubic Bezier,
                                         10 B-spline (NURBS) curve, 11 composite path,
                                         15 rail; 3D 4 sphere, 5 plane, 12 cylinder,
                                         13 line, 14 B-spline (NURBS) surface, 16 rail;
                                         NULL = all free (the group's "entities",
                                         src/bindings.cpp:412-446) */
    const int64_t* entity_off;        /* [ndof] start of the DOF's entity record in
                                         entity_data (ignored for tag 0); DOFs on one entity
                                         may share a record */
    const double* entity_data;        /* [entity_len] the records: value 0 the tag, then
                                         the kind's fields and payloads (egg_amd/csrc/entity.hpp
                                         lists them; src/geometry.hpp:1717-2520) */
    int64_t entity_len;
    int32_t curv_k;                   /* block-interface C2 curvature windows per DOF (2D; 0 =
                                         none; src/curvature.hpp, src/sweep.hpp:1461-1487) */
    const int32_t* curv_nodes;        /* [ndof*curv_k*4] global node ids of each window */
    const int32_t* curv_slot;         /* [ndof*curv_k] the DOF's place (0..3) in it, -1 = pad */
    const double* curv_weight;        /* [ndof*curv_k] window weight */
    int32_t dir_k;                    /* directional soft-energy samples per DOF (0 = none;
                                         src/directional.hpp, src/sweep.hpp:1488-1510, :1746-1760,
                                         wire src/bindings.cpp:393-409) */
    double dir_eps;                   /* their epsilon (x / sqrt(|x|^2 + eps^2)); one per session */
    const int32_t* dir_nodes;         /* [ndof*dir_k*4] global node ids, -1 pads */
    const int32_t* dir_slot;          /* [ndof*dir_k] the DOF's slot in the sample, -1 = pad */
    const int32_t* dir_kind;          /* [ndof*dir_k] 0 parallel, 1 line, 2 stem; others add nothing */
    const double* dir_ref;            /* [ndof*dir_k*D] frozen reference vector */
    const double* dir_axis;           /* [ndof*dir_k*D] projector axis of the 3D-only kinds (unused) */
    const double* dir_weight;         /* [ndof*dir_k] */
} eggmi_group;

/* The energy stencil (every (cell, corner) sample). */
typedef struct {
    int64_t num_samples;
    const int32_t* gc;
    const int32_t* gn[3];
    const double* s[3];
    const double* W_inv;
    int64_t w_rows; /* 1 or num_samples */
    const double* weight; /* [num_samples] energy weight, NULL = 1 (src/sweep.hpp:122, :301-302) */
    /* whole-sample directional table added to every reported energy and to
       the FAS line-search energies (src/sweep.hpp:1921, src/fas.hpp:233,
       wire src/bindings.cpp:471-480); 0 = none */
    int64_t dir_num;
    double dir_eps;
    const int32_t* dir_nodes;  /* [dir_num*4] global node ids, -1 pads */
    const int32_t* dir_kind;   /* [dir_num] */
    const double* dir_ref;     /* [dir_num*D] */
    const double* dir_axis;    /* [dir_num*D] (unused) */
    const double* dir_weight;  /* [dir_num] */
} eggmi_stencil;

/* Halo-padded structured tables (reference egg/smoothing/cpp_backend.py:
 * 679-691 structured_arrays). */
typedef struct {
    int32_t num_blocks;
    const int32_t* interior_shapes;          /* [num_blocks*dim] node counts */
    const int32_t* const* block_global_dof;  /* [num_blocks] -> [prod(shape)] row-major */
    int64_t n_halo;
    const int32_t* halo_src_block;   /* [n_halo] */
    const int32_t* halo_src_padded;  /* [n_halo*dim] */
    const int32_t* halo_dst_block;
    const int32_t* halo_dst_padded;
    int64_t n_sing;
    const int32_t* sing_block;   /* [n_sing] */
    const int32_t* sing_logical; /* [n_sing*dim] */
} eggmi_structured;

/* Build a device-resident session: validate the wire
 * (src/bindings.cpp:98-253), re-home it onto the halo-padded store
 * (src/bindings.cpp:758-1101) and upload once.
 * device: 0 = "auto", 1 = "gpu", 2 = "cpu" (no CPU backend: EGGMI_NO_DEVICE),
 * 3 = tables only (validation + remap on the host, no device state: only
 * eggmi_session_info / eggmi_session_tables / eggmi_mg_levels may follow).
 * Replaces make_structured_session (src/bindings.cpp:1153-1259). */
eggmi_status eggmi_session_create(int dim, const eggmi_group* groups, int n_groups, const eggmi_stencil* stencil,
                                  const eggmi_structured* structured, const double* X, int64_t num_nodes,
                                  int device, eggmi_session** out);

eggmi_status eggmi_session_destroy(eggmi_session* s);

/* n_sweeps of block-Jacobi; writes ceil(n/k) (k = report_every <= 0 ? n : k)
 * (energy, min det) reports. Replaces CppStructuredSweepSession::run
 * (src/bindings.cpp:1293-1307) -> StructuredExecutorT::run_jacobi
 * (src/structured_sweep.hpp:87-104). */
eggmi_status eggmi_run(eggmi_session* s, int n_sweeps, int phase, double delta, double omega, int report_every,
                       double* energies, double* mindets, int64_t capacity, int64_t* n_reports);

/* FAS V-cycles. Replaces CppStructuredSweepSession::run_fas
 * (src/bindings.cpp:1311-1335) -> StructuredExecutorT::run_fas
 * (src/structured_sweep.hpp:114-149) -> egg::run_fas (src/fas.hpp:398-602). */
eggmi_status eggmi_run_fas(eggmi_session* s, int n_cycles, int nu_pre, int nu_post, int nu_coarse, double omega,
                           int max_levels, int phase, double* energies, double* mindets, int64_t capacity,
                           int64_t* n_reports);

/* Coarse hierarchy shapes (src/bindings.cpp:1340-1356): shapes_out receives
 * level-major, block-major dim-tuples for levels 1..n (fine level
 * excluded); *n_levels the number of coarse levels, *n_blocks per level. */
eggmi_status eggmi_mg_levels(eggmi_session* s, int max_levels, int32_t* shapes_out, int64_t capacity,
                             int32_t* n_levels, int32_t* n_blocks);

/* X gathered back to global node order (src/bindings.cpp:1359-1376). */
/* Introspection of FAS coarse level `level` (1 = the first below the fine
 * grid; no reference counterpart, the reference keeps MgLevel internal,
 * src/mg_level.hpp:73-112): its slot count, its copy table (face-ghost
 * fills first, n_halo of them, then n_share owner -> copy broadcasts; node
 * slots of the level's padded layout) and its relaxed-DOF masks (free:
 * every DOF its sweeps relax; free_interior: the block-interior ones).
 * Level 0 is the fine store: its copy table is halo | fan | share, with the
 * fan mirrors counted in n_halo, and free_interior its synthesized DOFs.
 * Call once with null arrays for the sizes. Available in tables-only
 * sessions (device 3). */
eggmi_status eggmi_mg_level_tables(eggmi_session* s, int level, int64_t* n_slots, int64_t* n_halo,
                                   int64_t* n_share, int32_t* copy_src, int32_t* copy_dst, uint8_t* free_mask,
                                   uint8_t* free_interior);

/* The FAS transfer operators between coarse level `level` (>= 1) and its
 * parent applied to caller fields, for tests of their identities (the
 * reference tests its transfer kernels directly,
 * tests/cpp/test_multigrid_device.cpp:76-220, 389-470). Fields are
 * [dim][n_slots] doubles of the level they live on (n_slots from
 * eggmi_mg_level_tables; level 0 = the fine store, eggmi_session_info).
 * op 0: injection, in = parent field, out = child field (every interior
 * node); op 1: (1/4, 1/2, 1/4) full weighting of a ghost-inclusive parent
 * field, out = child field at its relaxed DOFs, zero elsewhere; op 2:
 * multilinear prolongation of a child field (the coarse change), out =
 * parent field at its relaxed nodes, zero at its other interior nodes;
 * op 3 (level 1, in unused): out = the tau-shifted coarse gradient at the
 * injected resident state, then the full-weighted fine gradient (two child
 * fields; equal by construction of tau, src/fas.hpp:522-532). */
eggmi_status eggmi_mg_probe(eggmi_session* s, int level, int op, const double* in, double* out);

/* Projection of n points q[n*dim] onto one entity and the orthonormal
 * tangent basis there, in FP64 with the projections the constrained sweep
 * uses (cold: no warm-start parameter): record[0..len-1] as in
 * group.entity_data (record[0] = 0: free, identity and the unit basis), dim
 * 2 or 3 the entity's. proj (n*dim) and basis (n*dim*k, column j of point i
 * at [(i*dim + r)*k + j], k the tangent dimension: 1 for curves, 2 for
 * surfaces, dim when free) may be null. Replaces the host geometry calls of
 * the binding: geometry_project / geometry_tangent / geometry_project_batch
 * (src/bindings.cpp:1872-1990, src/geometry.hpp:195-258, :2580-2612). */
eggmi_status eggmi_geometry_eval(int dim, const double* record, int64_t len, int64_t n, const double* q,
                                 double* proj, double* basis);

eggmi_status eggmi_get_X(eggmi_session* s, double* X_out, int64_t n);

/* Replace the resident positions (global order) — checkpoint resume. */
eggmi_status eggmi_set_X(eggmi_session* s, const double* X, int64_t n);

/* Session introspection for index-map parity (no reference counterpart;
 * the reference keeps these tables internal to StructuredRemap):
 * slot counts and the copy tables in node units (reference real offsets / dim). */
typedef struct {
    int64_t n_slots;      /* padded nodes + overflow */
    int64_t n_overflow;
    int64_t n_halo;       /* face-ghost entries */
    int64_t n_fan;        /* spare-ghost / overflow mirrors */
    int64_t n_share;      /* owner -> copy broadcasts */
    int64_t n_tiles;
    int64_t n_stored;     /* stored-stencil DOFs */
    int64_t n_synth;      /* synthesized interior DOFs */
    int32_t structured_energy; /* 1: energy stencil reduced from the block cells */
} eggmi_info;

eggmi_status eggmi_session_info(eggmi_session* s, eggmi_info* info);

/* copy tables (halo | fan | share order), g2s, block node bases */
eggmi_status eggmi_session_tables(eggmi_session* s, int32_t* copy_src, int32_t* copy_dst, int32_t* g2s,
                                  int64_t* block_base);

/* The packed device store, coordinate-major ([dim][n_slots], FP32). */
eggmi_status eggmi_get_packed(eggmi_session* s, float* out, int64_t n);

/* One fused halo + fan + share copy on the resident store
 * (src/structured_halo.hpp:219-237 fused_halo_broadcast). */
eggmi_status eggmi_halo_exchange(eggmi_session* s);

/* Select the sweep tile shape (autotuner hook); returns the variant count
 * in *n_variants when variant < 0. 2D: variants 0-3 put one lane on each of
 * a node's 4 cells (tiles of 32x2, 16x4, 64x1, 32x4 nodes), 4-5 one lane on
 * each of its 16 samples (16x1, 8x2 nodes), 6-7 one lane per cell without
 * LDS staging (64x1, 32x2), 8 and 11 two cells per lane (32x4, 64x2), 10
 * the same without LDS staging (32x4), 9 one lane per node (64x4), 12 and
 * 13 variant 0 with 4 / 8 lanes per stored-stencil DOF instead of 16. 3D:
 * one lane per cell, tiles of 8x4x2 and 4x4x4 nodes, 2 = 8x4x2 without LDS
 * staging, 3 = 8x4x2 with two cells per lane, 4 = variant 2 with 16 lanes
 * per stored-stencil DOF instead of 64. Only variants 0-3 (2D) and
 * 0-1 (3D) have a learned-proposal kernel; setting a proposal on a session
 * with another variant selects variant 0. */
eggmi_status eggmi_set_tile_variant(eggmi_session* s, int variant, int* n_variants);

/* Energy and min det A of the resident state (the report reduction of
 * eggmi_run, src/sweep.hpp:1881-1922, FP64 accumulation) without sweeping. */
eggmi_status eggmi_energy(eggmi_session* s, int phase, double delta, double* energy, double* mindet);

/* Autotuner (no reference counterpart): time n_sweeps plain sweeps (omega)
 * with every tile variant from the session's current state, which is
 * restored after each trial, and keep the fastest. The variants sum a patch
 * in the same order except the one-lane-per-sample ones, and differ at most
 * by FP32 rounding (multiply-add fusion). ms[v] (v < capacity) receives each
 * variant's time, *best the variant selected. */
eggmi_status eggmi_autotune(eggmi_session* s, int n_sweeps, double omega, double* ms, int capacity, int* best);

/* Tile variant of each FAS coarse level (no reference counterpart):
 * variants[i] for coarse level i + 1 (level 0 is the fine grid, which keeps
 * the session's variant), -1 = the session's variant; levels beyond n
 * follow the session's variant, and n = 0 clears every choice. Coarse levels
 * never run the learned proposal, so every variant is allowed there. A
 * value V + v (V = the tile variant count) smooths the level in one
 * single-workgroup launch per smoothing call, walking variant v's tiles in
 * turn with a workgroup barrier between them (same bits as v). */
eggmi_status eggmi_set_level_variants(eggmi_session* s, const int* variants, int n);

/* Level autotuner (no reference counterpart). Coarse levels have far fewer
 * DOFs than the chip has SIMDs, so a sweep there lasts about one wave's
 * lifetime and a different tile shape can win than on the fine grid. For
 * each of the first max_levels - 1 coarse levels, from the resident fine
 * state restricted down the chain (as a cycle's first restriction, without
 * smoothing), time n_sweeps FAS-shifted sweeps of `phase` (barrier phases
 * only), replayed as one captured hipGraph as a cycle replays them, with
 * every tile variant and keep the fastest; levels of at most 64 tiles are
 * also timed as one single-workgroup launch per variant (level variant
 * V + v). The fine state is not changed. ms[l * 2V + v] (< capacity; V =
 * the tile variant count) receives level l + 1's time with level variant
 * v (-1: not timed), best[l] (< best_capacity) its pick, *n_levels the
 * number of levels tuned. */
eggmi_status eggmi_autotune_levels(eggmi_session* s, int n_sweeps, double omega, int max_levels, int phase,
                                   double* ms, int64_t capacity, int* best, int best_capacity, int* n_levels);

/* Multi-sweep launches (no reference counterpart: the reference launches
 * every sweep, src/sweep.hpp:1938-1984): level `level` (0 = the fine grid,
 * i >= 1 = FAS coarse level i, which a run_fas must have prepared) runs k
 * block-Jacobi sweeps per launch (k <= 8) over core tiles of ty x tx nodes
 * with multi-sweep kernel variant `kernel` (< *n_kernels when n_kernels is
 * non-null; eggmi_multi_sweep_info). Each workgroup stages its tile's
 * dependency closure k rings deep in LDS and recomputes the outer rings with
 * the one-sweep kernel's per-DOF device code, so the result equals k
 * one-sweep launches up to a few float ulps (multiply-add fusion may
 * differ between the two kernels). k <= 1 returns the
 * level to one launch per sweep. 2D levels without special stored DOFs
 * (constrained, curvature or directional windows) or per-sample energy
 * weights, and closures within the LDS; INVALID_ARGUMENT otherwise. Runs
 * with the learned proposal keep one launch per sweep. */
eggmi_status eggmi_set_multi_sweep(eggmi_session* s, int level, int k, int ty, int tx, int kernel, int* n_kernels);

/* A level's multi-sweep tables: info[0..7] = sweeps per launch in use (0:
 * off), ty, tx, kernel variant, tiles, LDS nodes per buffer plane, core DOFs
 * and DOF updates per launch over its k sub-sweeps (their ratio to k x core
 * DOFs is the redundant recompute). With check_err non-null, also a
 * host-side check of the tables: the largest difference between k sweeps
 * of a synthetic nonlinear update of every DOF from its stencil (seeded
 * state) run through the tables' schedule and k plain synchronous sweeps
 * of it (0 when every closure, region, push and write is right). */
eggmi_status eggmi_multi_sweep_info(eggmi_session* s, int level, int64_t* info, double* check_err, unsigned seed);

/* Autotuner over multi-sweep candidates on the fine grid (no reference
 * counterpart): cand[4 i .. 4 i + 3] = (k, ty, tx, kernel) for i < n_cands;
 * times n_sweeps plain sweeps (omega) per candidate from the current state
 * (restored after each trial) and keeps the fastest if it beats base_ms
 * (else the previous setting stays). ms[i] receives each time (+inf: the
 * candidate's closures outgrow the LDS). */
eggmi_status eggmi_autotune_multi(eggmi_session* s, int n_sweeps, double omega, const int32_t* cand, int n_cands,
                                  double base_ms, double* ms);

/* The same per FAS coarse level (no reference counterpart): for each of the
 * first max_levels - 1 coarse levels, from the resident fine state
 * restricted down the chain, time one smoothing call of n_sweeps FAS-shifted
 * sweeps of `phase` (captured as a cycle replays it) with the level's
 * one-sweep setting and with each candidate, and keep the fastest.
 * ms[l * (n_cands + 1) + 0] (< capacity) receives level l + 1's one-sweep
 * time, [.. + 1 + i] candidate i's (-1: not buildable); *n_levels the
 * levels tuned. The fine state is not changed. */
eggmi_status eggmi_autotune_levels_multi(eggmi_session* s, int n_sweeps, double omega, int max_levels, int phase,
                                         const int32_t* cand, int n_cands, double* ms, int64_t capacity,
                                         int* n_levels);

/* Device time of the last eggmi_run / eggmi_run_fas call in ms (HIP events
 * on the session stream), the number of sweep-kernel launches it made and,
 * when kernel timing is on, the summed duration of those launches. */
eggmi_status eggmi_last_timing(eggmi_session* s, double* ms, int64_t* sweep_launches, double* sweep_kernel_ms);

/* Kernel nodes of the captured FAS cycle graph the last eggmi_run_fas call
 * replayed (0 when it ran uncaptured). A diagnostic with no counterpart in
 * the reference: the launch count of one V-cycle. */
eggmi_status eggmi_fas_cycle_kernels(eggmi_session* s, int64_t* kernels);

/* Bracket every sweep launch with HIP events (measurement only). */
eggmi_status eggmi_set_kernel_timing(eggmi_session* s, int on);

/* Replay plain sweeps in chunks of `chunk` (even; 0 = off, the default)
 * captured once as a hipGraph per starting buffer, instead of one launch
 * each, and (chunk > 0) replay each FAS cycle as one captured graph with its
 * line-search ladder evaluated and decided on the device (no reference
 * counterpart; the same bits as single launches). Reports still come at the
 * requested cadence; per-launch timing and counters fall back to single
 * launches. */
eggmi_status eggmi_set_graph_chunk(eggmi_session* s, int chunk);

/* Diagnostic counters (measurement only; no reference counterpart): when on,
 * every sweep launch counts the line-search trials it evaluated and the DOF
 * updates it made, per multigrid level. */
eggmi_status eggmi_set_counters(eggmi_session* s, int on);

/* Counters of the last run / run_fas: out[3*l] line-search trials, out[3*l+1]
 * DOF updates and out[3*l+2] accepted learned proposals of level l (0 = fine),
 * for l < n_levels; zeros when counting was off. */
eggmi_status eggmi_last_counters(eggmi_session* s, int64_t* out, int n_levels);

/* ---- Learned Newton-step proposal (no reference counterpart). A small MLP
 * maps a synthesized interior DOF's one-ring (local frame, scaled by the
 * local edge length) and its previous displacement to a proposed
 * displacement; eggmi_run accepts a proposal only if the DOF's patch energy
 * drops and every incident det stays > 0, and runs the plain Newton step
 * otherwise. Inference and training use the BF16 matrix cores with FP32
 * accumulation; the FP32 master parameters are laid out
 * W1[n_in][hidden], b1[hidden], W2[hidden][hidden], b2[hidden],
 * W3[hidden][n_out], b3[n_out]. Features are rows of n_in_padded floats. */
typedef struct eggmi_mlp eggmi_mlp;

/* A network for dim 2 (22 -> 64 -> 64 -> 2) or 3 (85 -> 128 -> 128 -> 3),
 * Glorot-initialised from `seed`, with room for `capacity` training pairs
 * on the device. */
eggmi_status eggmi_mlp_create(int dim, uint64_t seed, int64_t capacity, eggmi_mlp** out);
eggmi_status eggmi_mlp_destroy(eggmi_mlp* m);
/* shape[5] = {dim, n_in, n_in_padded, hidden, n_out}; any output may be null. */
eggmi_status eggmi_mlp_shape(eggmi_mlp* m, int32_t* shape, int64_t* n_params, int64_t* n_data, int64_t* capacity);
eggmi_status eggmi_mlp_get_params(eggmi_mlp* m, float* out, int64_t n);
/* Replace the parameters (resets the optimiser state). */
eggmi_status eggmi_mlp_set_params(eggmi_mlp* m, const float* in, int64_t n);
/* Y[rows][n_out] for features X[rows][n_in_padded] (rounded to BF16). */
eggmi_status eggmi_mlp_forward(eggmi_mlp* m, const float* X, int64_t rows, float* Y);
/* Append training pairs X[rows][n_in_padded], T[rows][n_out]. */
eggmi_status eggmi_mlp_data_add(eggmi_mlp* m, const float* X, const float* T, int64_t rows);
/* Copy rows [offset, offset+rows) of the training set out (X as stored, BF16). */
eggmi_status eggmi_mlp_data_get(eggmi_mlp* m, float* X, float* T, int64_t offset, int64_t rows);
eggmi_status eggmi_mlp_data_clear(eggmi_mlp* m);
/* Gradient of the mean squared error over the training rows rows_idx[batch]
 * (batch a positive multiple of 256); *loss = that mean squared error. */
eggmi_status eggmi_mlp_gradient(eggmi_mlp* m, const int32_t* rows_idx, int64_t batch, float* grad, double* loss);
/* Adam (lr, betas 0.9 / 0.999) over shuffled batches of the training set;
 * epoch_loss[epochs] receives each epoch's mean squared error. */
eggmi_status eggmi_mlp_train(eggmi_mlp* m, int epochs, int64_t batch, float lr, uint64_t seed, double* epoch_loss);
/* Run n_snap x (every + horizon) plain sweeps (omega) on the session and
 * append, at each snapshot, one pair per synthesized interior DOF: its
 * features and its displacement over the next `horizon` sweeps (local frame,
 * over h). *added = pairs appended (stops early when the set is full). */
eggmi_status eggmi_session_harvest(eggmi_session* s, eggmi_mlp* m, int n_snap, int every, int horizon,
                                   double omega, int64_t* added);
/* Use `m` (null: none) as the proposal of subsequent eggmi_run calls in the
 * barrier and shape_size phases and of eggmi_run_fas's fine-level smoothing
 * sweeps, its displacement scaled by `scale`. mode 0:
 * an accepted proposal is the DOF's update; mode 1 (predict-correct): the
 * plain Newton update (halvings, omega blend) continues from an accepted
 * proposal. The session borrows m: keep it alive (or unset it) until the
 * session is destroyed. */
eggmi_status eggmi_session_set_proposal(eggmi_session* s, eggmi_mlp* m, int mode, float scale);
/* Proposal cadence (default 1): eggmi_run proposes in sweeps 0, every,
 * 2 every, ... of each call (the other sweeps run the plain kernel of the
 * session's tile variant); with every > 1, each fine-level smoothing call of
 * eggmi_run_fas proposes in its first sweep only. No reference counterpart:
 * the learned proposal is an addition of this library's. */
eggmi_status eggmi_session_set_proposal_every(eggmi_session* s, int every);

/* Jacobi-preconditioned conjugate gradients in FP64 on a CSR operator
 * (rowptr[n+1], col/val[nnz]; SPD expected): the PCG skeleton of the
 * reference's control-net Gauss-Newton step, pcg_solve
 * (src/control_solve.hpp:761-818): x0 = 0, stop when |r| <= rtol |b|,
 * stop on non-positive curvature. recurrence 0: the reference's (two
 * launches per iteration); 1: Chronopoulos-Gear, the same iterates in exact
 * arithmetic with the three inner products of an iteration reduced together
 * (one launch per iteration). x receives the solution; *status is 1
 * converged, 2 non-positive curvature, 3 zero right-hand side, 0 max_iter
 * reached; *solve_ms (may be null) the device time of the solve, uploads
 * excluded. */
eggmi_status eggmi_pcg_csr(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                           const double* b, double* x, double rtol, int max_iter, int recurrence, int* iters,
                           double* relres, int* status, double* solve_ms);

/* The same solver on the matrix-free (2 dim + 1)-point lattice operator of
 * shape[0] x .. x shape[dim-1] unknowns (row-major, last axis fastest):
 * `diag` on the diagonal, `off` for each neighbour one step along an axis
 * inside the lattice (Dirichlet closure). diag = 2 dim, off = -1 is the
 * lattice graph Laplacian with Dirichlet boundary. No counterpart in the
 * reference's binding: it is the benchmark's spacing system without the
 * CSR index traffic, for the PCG of src/control_solve.hpp:761-818. Outputs
 * and recurrence as for eggmi_pcg_csr. */
eggmi_status eggmi_pcg_lattice(int dim, const int32_t* shape, double diag, double off, const double* b, double* x,
                               double rtol, int max_iter, int recurrence, int* iters, double* relres, int* status,
                               double* solve_ms);

/* eggmi_pcg_csr without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight solver loops,
 * harnesses that forbid hipMalloc). Every pointer is DEVICE memory (rowptr
 * int64[n+1], col int32[nnz], val, b, x double); the operator is not
 * validated (eggmi_pcg_csr's checks need a host copy). The work area is a
 * file-scope device array of the library (one per GPU, ~18 MiB, held until
 * the process ends): n <= 262144 (EGGMI_INVALID_ARGUMENT beyond). All calls
 * on one GPU share that array, so calls that overlap on one GPU race.
 * stream: a hipStream_t, NULL for the default stream; the call returns when
 * the solve is done. Same kernels, iterates and outputs as eggmi_pcg_csr. */
eggmi_status eggmi_pcg_csr_noalloc(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                                   const double* b, double* x, double rtol, int max_iter, int recurrence,
                                   void* stream, int* iters, double* relres, int* status);

/* ---- control-net reduced Gauss-Newton (one block) */

/* One block's control net (the reference's `control` wire,
 * egg/smoothing/control_backend.py:124-136, unpacked at src/bindings.cpp
 * into ControlNetHostT): the block's nodes are X = M C + b with M the tensor
 * product of per-axis banded windows. Node i of axis k reads controls
 * off[k][i] .. off[k][i] + k_supp - 1 with weights wt[k][i * k_supp ..].
 * C0: the lattice (row-major over ctrl_shape, last axis fastest, dim values
 * per control); free_idx: flat lattice indices of the free controls; b:
 * the block's Boolean-sum offset in its node order (row-major over the
 * block's node shape, dim values per node), NULL = 0. Sliding wall controls
 * (the wire's walls / X0) have no counterpart here. */
typedef struct {
    int32_t block;
    int32_t ctrl_shape[3];
    int32_t k_supp;
    const int32_t* off[3];
    const double* wt[3];
    const double* C0;
    int64_t n_free;
    const int32_t* free_idx;
    const double* b;
} eggmi_control;

/* Attach a control net to a session and evaluate it into the resident
 * state (the block's nodes become M C0 + b). Replaces the session
 * constructor's `control=` argument (src/bindings.cpp:1153-1259,
 * ControlSessionT ctor src/control_solve.hpp:352-480). */
eggmi_status eggmi_session_set_control(eggmi_session* s, const eggmi_control* c);
/* Upload a full lattice (n = n_ctrl * dim doubles) and re-evaluate the net
 * into the resident state; download it. Replace set_C / get_C
 * (src/control_solve.hpp:493-509). */
eggmi_status eggmi_control_set_C(eggmi_session* s, const double* C, int64_t n);
eggmi_status eggmi_control_get_C(eggmi_session* s, double* C, int64_t n);

/* ---- control-net reduced Gauss-Newton over several blocks */

/* Several blocks' nets over one stacked control vector (the reference's
 * multi-block control wire, selected by its "blocks" key and unpacked at
 * src/bindings.cpp:593-740 into ControlTopoHostT; session
 * ControlTopoSessionT, src/control_solve.hpp:1585-2837). nets[i] gives
 * net i's block, ctrl_shape, k_supp, off, wt and b with the meaning they have
 * for one block; its C0, n_free and free_idx are not read (a net's freedom
 * lives in R); at
 * most one net per block. C0: the nets' lattices stacked in net order
 * (n_stack * dim doubles, n_stack the total control count). The
 * reduction: dC = R dq with R an (n_stack * dim) x nq CSR over the stacked
 * components (row r = control * dim + component; r_off has n_stack * dim
 * + 1 entries); rows without entries keep their component at C0. Optional
 * homogeneous penalty rows P (n_pen x n_stack * dim CSR, p_off n_pen + 1
 * entries, weights p_w): energy 1/2 sum_j p_w[j] (P_j C)^2, composed into
 * the Gauss-Newton system and the line-search energy. Every block node
 * takes the value of its global point's owner (the node at the point's
 * root slot); blocks without a net keep their nodes, except where they
 * copy a net block's owner. Sliding-wall frames (the wire's slide,
 * wall_faces, seam_face) have no counterpart here. */
typedef struct {
    int32_t n_nets;
    const eggmi_control* nets;
    const double* C0;
    int64_t nq;
    const int32_t* r_off;
    const int32_t* r_col;
    const double* r_coef;
    int64_t n_pen;
    const int32_t* p_off;
    const int32_t* p_col;
    const double* p_coef;
    const double* p_w;
} eggmi_control_topo;

/* Attach the nets (replacing any control net of the session) and evaluate
 * them into the resident state with the owner-to-copy refresh
 * (ControlTopoSessionT ctor + eval_to_x, src/control_solve.hpp:1590-2055,
 * :2187-2191). eggmi_control_set_C / get_C / eggmi_run_control then act on
 * the stacked vector (n_stack * dim doubles). */
eggmi_status eggmi_session_set_control_topo(eggmi_session* s, const eggmi_control_topo* t);
/* Replace the reduction (set_reduction, src/control_solve.hpp:2057-2117;
 * the reference binding set_control_reduction, src/bindings.cpp:1475-1494):
 * same row space, any column count nq. */
eggmi_status eggmi_control_set_reduction(eggmi_session* s, int64_t nq, const int32_t* r_off, const int32_t* r_col,
                                         const double* r_coef);
/* Replace the penalty rows (set_penalty, src/control_solve.hpp:2120-2169;
 * binding set_control_penalty, src/bindings.cpp:1496-1516); n_pen = 0
 * removes them. */
eggmi_status eggmi_control_set_penalty(eggmi_session* s, int64_t n_pen, const int32_t* p_off, const int32_t* p_col,
                                       const double* p_coef, const double* p_w);
/* Replace net `net`'s offset b (n = its block's node count * dim doubles)
 * and re-evaluate the nets into the state (set_b, src/control_solve.hpp:
 * 2172-2184; binding set_control_b, src/bindings.cpp:1518-1531). */
eggmi_status eggmi_control_set_b(eggmi_session* s, int32_t net, const double* b, int64_t n);

/* Options of the reduced Gauss-Newton loop (ControlOptions,
 * src/control_solve.hpp:308-329). */
typedef struct {
    int32_t n_outer;       /* outer iterations at most */
    double grad_tol;       /* converged when max |G| < grad_tol */
    double alpha_min;      /* line-search floor */
    double lm0;            /* initial Levenberg damping */
    int32_t lm_tries;      /* damping ladder per outer iteration */
    int32_t pcg_max_iter;
    double pcg_rtol;       /* PCG relative-residual floor */
    int32_t pcg_forcing;   /* 1: tolerance max(pcg_rtol, min(0.5, sqrt(max|G| / max|G|_0))) */
} eggmi_control_options;

/* Run the loop in place on the session's state (ControlSessionT::run_plain,
 * src/control_solve.hpp:533-676; with several blocks' nets attached
 * ControlTopoSessionT::run, :2209-2425, over the reduced DOFs q: gradient
 * R^T (M^T g + P^T W P C), the same PCG on R^T (A + P^T W P) R + lam I
 * preconditioned by the diagonal of R^T Pj R with Pj the per-control Jacobi
 * blocks, the penalty energy added in the line search): reduced gradient, Jacobi-preconditioned
 * matrix-free FP64 PCG on (A + lam I) dC = -G, line search of M dC on the
 * global fine energy (FP64 reductions; accept iff finite, E <= E0 + 1e-6 and
 * min det > 0), damping x10 (floored at 1e-8 trace(P) / n) on rejection,
 * lam / 3 after a full step. phase: barrier or shape_size (untangle is
 * EGGMI_INVALID_ARGUMENT). energies / mindets receive the initial state's
 * and every accepted iterate's (capacity >= n_outer + 1), alphas the
 * accepted steps, pcg_iters / pcg_relres one entry per PCG solve (capacity
 * pcg_capacity >= n_outer * lm_tries); *n_reports, *n_pcg, *iters and
 * *converged the counts and flags. */
eggmi_status eggmi_run_control(eggmi_session* s, const eggmi_control_options* o, int phase, double* energies,
                               double* mindets, double* alphas, int64_t capacity, int64_t* n_reports,
                               double* pcg_iters, double* pcg_relres, int64_t pcg_capacity, int64_t* n_pcg,
                               int32_t* iters, int32_t* converged);

const char* eggmi_last_error(void);

/* Library liveness (reference cpp_core.ping, src/bindings.cpp:1544). */
int eggmi_ping(void);

#ifdef __cplusplus
}
#endif

#endif /* EGGMI_H */
