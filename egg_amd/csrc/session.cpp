// Host side of the MI355X structured smoother (see session.hpp).
#include "session.hpp"
#include "directional.hpp"
#include "entity.hpp"
#include "mlp_host.hpp"

#include <limits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <utility>

namespace eggmi
{

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Internal(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ layout

void Layout::build(int d, const std::vector<std::array<int, 3>>& shapes)
{
    dim = d;
    shape = shapes;
    base.assign(shapes.size(), 0);
    stride.assign(shapes.size(), {0, 0, 0});
    long long acc = 0;
    for (size_t b = 0; b < shapes.size(); ++b) {
        base[b] = acc;
        long long st = 1;
        for (int k = d - 1; k >= 0; --k) {
            stride[b][k] = st;
            st *= shapes[b][k] + 2;
        }
        acc += st;
    }
    padded_nodes = acc;
    overflow = 0;
}

long long Layout::padded_slot(int b, const int* padded) const
{
    long long s = base[b];
    for (int k = 0; k < dim; ++k) s += (long long)padded[k] * stride[b][k];
    return s;
}

long long Layout::interior_slot(int b, const int* logical) const
{
    long long s = base[b];
    for (int k = 0; k < dim; ++k) s += (long long)(logical[k] + 1) * stride[b][k];
    return s;
}

long long Layout::block_padded_count(int b) const
{
    long long n = 1;
    for (int k = 0; k < dim; ++k) n *= shape[b][k] + 2;
    return n;
}

// ------------------------------------------------------------------ level

void Level::make_tiles(int v)
{
    const TileShape ts = tile_variant_shape(lay.dim, v);
    variant = v;
    const int G = tile_variant_stored_lanes(lay.dim, v);
    if (G != st_lanes) upload_stored(G);
    // axis 0 <- tz (3D) / ty (2D), last axis <- tx
    const int step[3] = {lay.dim == 2 ? ts.ty : ts.tz, lay.dim == 2 ? ts.tx : ts.ty, lay.dim == 2 ? 1 : ts.tx};
    tiles.clear();
    for (size_t b = 0; b < lay.shape.size(); ++b) {
        const auto& n = lay.shape[b];
        // bounding box of the block's synthesized DOFs (logical, interior slots only)
        int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
        const int n2 = lay.dim == 3 ? n[2] : 1;
        for (int i = 0; i < n[0]; ++i)
            for (int j = 0; j < n[1]; ++j)
                for (int k = 0; k < n2; ++k) {
                    const int lg[3] = {i, j, k};
                    if (!mask[lay.interior_slot((int)b, lg)]) continue;
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], lg[a]);
                        hi[a] = std::max(hi[a], lg[a]);
                    }
                }
        if (hi[0] < 0) continue;
        for (int i = lo[0]; i <= hi[0]; i += step[0])
            for (int j = lo[1]; j <= hi[1]; j += step[1])
                for (int k = lo[2]; k <= hi[2]; k += step[2]) {
                    TileDesc t {};
                    t.b = block_desc(b);
                    t.blk = (int)b;
                    t.i0 = i;
                    t.j0 = j;
                    t.k0 = k;
                    tiles.push_back(t);
                }
    }
    d_tiles.upload(tiles);
}

void Level::build_push()
{
    // root source of a copy: follow owner -> copy broadcasts backwards (a
    // table entry whose source is itself a broadcast destination)
    const long long ns = lay.n_slots();
    std::vector<int> share_from(ns, -1);
    for (size_t q = copy_src.size() - n_share; q < copy_src.size(); ++q) share_from[copy_dst[q]] = copy_src[q];
    std::vector<uint8_t> is_dof(mask.begin(), mask.end());
    for (int d : st_dof) is_dof[d] = 1;
    std::vector<std::pair<int, int>> all;  // (root, destination) of every copy
    is_dst.assign(ns, 0);
    for (size_t q = 0; q < copy_src.size(); ++q) {
        int r = copy_src[q];
        for (int guard = 0; share_from[r] >= 0 && guard < 8; ++guard) r = share_from[r];
        all.emplace_back(r, copy_dst[q]);
        is_dst[copy_dst[q]] = 1;
    }
    auto csr = [ns](const std::vector<std::pair<int, int>>& pairs, std::vector<int>& off, std::vector<int>& dst) {
        off.assign(ns + 1, 0);
        for (auto& pr : pairs) ++off[pr.first + 1];
        for (long long i = 0; i < ns; ++i) off[i + 1] += off[i];
        dst.assign(pairs.size(), 0);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (auto& pr : pairs) dst[fill[pr.first]++] = pr.second;
    };
    std::vector<std::pair<int, int>> pushes;
    for (auto& pr : all)
        if (is_dof[pr.first]) pushes.push_back(pr);
    csr(pushes, cp_off, cp_dst);
    n_push = (long long)pushes.size();
    csr(all, rp_off, rp_dst);
}

BlockDesc Level::block_desc(size_t b) const
{
    BlockDesc d {};
    d.base = lay.base[b];
    for (int k = 0; k < 3; ++k) {
        d.n[k] = k < lay.dim ? lay.shape[b][k] : 1;
        d.st[k] = k < lay.dim ? lay.stride[b][k] : 0;
    }
    d.pad = 0;
    return d;
}

void Level::upload_tables()
{
    std::vector<BlockDesc> bd(lay.shape.size());
    for (size_t b = 0; b < bd.size(); ++b) bd[b] = block_desc(b);
    blocks.upload(bd);
    d_copy_src.upload(copy_src);
    d_copy_dst.upload(copy_dst);
    d_cp_off.upload(cp_off);
    d_cp_dst.upload(cp_dst);
    d_rp_off.upload(rp_off);
    d_rp_dst.upload(rp_dst);
    d_is_dst.upload(is_dst);
    d_mask.upload(mask);
    d_st_dof.upload(st_dof);
    d_con_tag.upload(con_tag);
    d_con_off.upload(con_off);
    d_con_blob.upload(con_blob);
    d_con_epoch.upload(std::vector<int>(con_tag.size(), -1));
    d_run_epoch.upload(std::vector<int>(1, run_epoch));
    reset_seeds();
    d_curv_nodes.upload(curv_nodes);
    d_curv_slot.upload(curv_slot);
    d_curv_w.upload(curv_w);
    d_dir_nodes.upload(dir_nodes);
    d_dir_slot.upload(dir_slot);
    d_dir_kind.upload(dir_kind);
    d_dir_ref.upload(dir_ref);
    d_dir_w.upload(dir_w);
    st_lanes = 0;  // the stored layout follows the tile variant (make_tiles)
    // structured energy cells
    cell_off.assign(lay.shape.size() + 1, 0);
    for (size_t b = 0; b < lay.shape.size(); ++b) {
        long long c = 1;
        for (int k = 0; k < lay.dim; ++k) c *= std::max(0, lay.shape[b][k] - 1);
        cell_off[b + 1] = cell_off[b] + c;
    }
    for (size_t b = 0; b + 1 < cell_off.size(); ++b)
        if (cell_off[b + 1] - cell_off[b] >= (1ll << 31)) throw InvalidArgument("a block has 2^31 or more cells");
    ncells = cell_off.back();
    d_cell_off.upload(cell_off);
}

// Stored occurrences in the padded lane layout of StoredView for groups of G
// lanes per DOF: occurrence p of DOF d goes to round p / G, lane p % G.
void Level::upload_stored(int G)
{
    const int D = lay.dim;
    const long long n = (long long)st_dof.size(), total = (long long)st_gc.size();
    int rounds = 1;
    for (long long d = 0; d < n; ++d) rounds = std::max(rounds, (st_off[d + 1] - st_off[d] + G - 1) / G);
    st_rounds = rounds;
    st_padded = n * rounds * G;
    std::vector<int> gc(st_padded, 0), gn((size_t)D * st_padded, 0);
    std::vector<uint8_t> meta(st_padded, kStoredPad);
    std::vector<float> W(st_W.empty() ? 0 : (size_t)D * D * st_padded, 0.0f);
    std::vector<float> wt(st_wt.empty() ? 0 : (size_t)st_padded, 0.0f);
    for (long long d = 0; d < n; ++d)
        for (int p = st_off[d]; p < st_off[d + 1]; ++p) {
            const int o = p - st_off[d];
            const long long q = (d * rounds + o / G) * G + o % G;
            gc[q] = st_gc[p];
            meta[q] = st_meta[p];
            for (int k = 0; k < D; ++k) gn[(size_t)k * st_padded + q] = st_gn[(size_t)k * total + p];
            for (int e = 0; e < D * D && !W.empty(); ++e) W[(size_t)D * D * q + e] = st_W[(size_t)D * D * p + e];
            if (!wt.empty()) wt[q] = st_wt[p];
        }
    d_st_gc.upload(gc);
    d_st_gn.upload(gn);
    d_st_meta.upload(meta);
    d_st_W.upload(W);
    d_st_wt.upload(wt);
    st_lanes = G;
}

void Level::reset_seeds()
{
    d_con_seed.upload(std::vector<float>(2 * con_tag.size(), std::numeric_limits<float>::quiet_NaN()));
}

Level::ConState Level::save_con_state() const
{
    ConState c;
    c.run_epoch = run_epoch;
    c.seed.resize(d_con_seed.n);
    c.epoch.resize(d_con_epoch.n);
    if (!c.seed.empty())
        EGGMI_HIP(hipMemcpy(c.seed.data(), d_con_seed.p, c.seed.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (!c.epoch.empty())
        EGGMI_HIP(hipMemcpy(c.epoch.data(), d_con_epoch.p, c.epoch.size() * sizeof(int), hipMemcpyDeviceToHost));
    return c;
}

void Level::restore_con_state(const ConState& c)
{
    if (con_tag.empty()) return;
    run_epoch = c.run_epoch;
    d_con_seed.upload(c.seed);
    d_con_epoch.upload(c.epoch);
    d_run_epoch.upload(std::vector<int>(1, run_epoch));
}

void Level::next_run(hipStream_t s)
{
    if (con_tag.empty()) return;
    ++run_epoch;
    launch_set_int(d_run_epoch.p, run_epoch, s);
}

SweepArgs Level::sweep_args(const float* xin, float* xout, float omega, float delta, const float* tau) const
{
    SweepArgs a;
    a.xin = xin;
    a.xout = xout;
    a.plane = plane;
    a.tiles = d_tiles.p;
    a.n_tiles = (int)tiles.size();
    a.blocks = blocks.p;
    a.mask = d_mask.p;
    a.st.n = (int)(st_dof.size() - con_tag.size());  // the constrained DOFs follow (con_*)
    a.con_n = (int)con_tag.size();
    a.con_tag = d_con_tag.p;
    a.con_blob = d_con_blob.p;
    a.con_off = d_con_off.p;
    a.con_seed = d_con_seed.p;
    a.con_epoch = d_con_epoch.p;
    a.run_epoch = d_run_epoch.p;
    a.curv_k = con_tag.empty() ? 0 : curv_k;
    a.curv_nodes = d_curv_nodes.p;
    a.curv_slot = d_curv_slot.p;
    a.curv_w = d_curv_w.p;
    if (!con_tag.empty() && dir_k > 0)
        a.dir = DirView{dir_k, dir_eps, d_dir_nodes.p, d_dir_slot.p, d_dir_kind.p, d_dir_ref.p, d_dir_w.p};
    a.st.rounds = st_rounds;
    a.st.dof = d_st_dof.p;
    a.st.total = st_padded;
    a.st.gc = d_st_gc.p;
    a.st.gn = d_st_gn.p;
    a.st.meta = d_st_meta.p;
    a.st.W = st_W.empty() ? nullptr : d_st_W.p;
    a.st.wt = st_wt.empty() ? nullptr : d_st_wt.p;
    a.omega = omega;
    a.delta = delta;
    a.tau = tau;
    a.cp_off = n_push ? d_cp_off.p : nullptr;
    a.cp_dst = n_push ? d_cp_dst.p : nullptr;
    for (int i = 0; i < 9; ++i) a.Wu[i] = Wu[i];
    return a;
}

// ------------------------------------------------------------------ validation

namespace
{

void check_index_range(const int32_t* v, long long n, long long N, const std::string& grp, const char* name)
{
    for (long long i = 0; i < n; ++i) {
        if (v[i] < 0 || v[i] >= N) {
            throw InvalidArgument("validate_context: " + grp + " " + name + " index " + std::to_string(v[i]) +
                                  " out of range [0, " + std::to_string(N) + ")");
        }
    }
}

// Mirrors the checks of the reference's validate_context (src/bindings.cpp:98-253)
// for the fields this path consumes.
void validate_dir(int dim, const int32_t* nodes, const int32_t* slot, const int32_t* kind, const double* ref,
                  const double* wt, double eps, long long n, long long N, const std::string& grp);

void validate(int dim, const eggmi_group* groups, int n_groups, const eggmi_stencil* es, long long N)
{
    const int dd = dim * dim;
    for (int gi = 0; gi < n_groups; ++gi) {
        const eggmi_group& g = groups[gi];
        const std::string grp = "group[" + std::to_string(gi) + "]";
        if (g.ndof < 0 || g.total_samples < 0) throw InvalidArgument("validate_context: " + grp + " negative size");
        if (g.w_rows != 1 && g.w_rows != g.total_samples)
            throw InvalidArgument("validate_context: " + grp + " W_inv size " + std::to_string(g.w_rows * dd) +
                                  " != " + std::to_string(g.total_samples * dd));
        for (long long p = 0; p < g.total_samples; ++p) {
            if (g.role[p] < -1 || g.role[p] > dim)
                throw InvalidArgument("validate_context: " + grp + " role " + std::to_string(g.role[p]) +
                                      " out of range [-1, " + std::to_string(dim) + "]");
        }
        long long sumP = 0;
        for (long long d = 0; d < g.ndof; ++d) {
            if (g.P_of[d] < 0) throw InvalidArgument("validate_context: " + grp + " negative P_of");
            sumP += g.P_of[d];
        }
        if (sumP != g.total_samples)
            throw InvalidArgument("validate_context: " + grp + " sum(P_of) " + std::to_string(sumP) +
                                  " != total_samples " + std::to_string(g.total_samples));
        check_index_range(g.gc, g.total_samples, N, grp, "gc");
        for (int k = 0; k < dim; ++k) {
            const std::string nm = "gn" + std::to_string(k);
            check_index_range(g.gn[k], g.total_samples, N, grp, nm.c_str());
        }
        check_index_range(g.dof_idx, g.ndof, N, grp, "dof_idx");
        for (int k = 0; k < dim; ++k)
            for (long long p = 0; p < g.total_samples; ++p)
                if (g.s[k][p] != 1.0 && g.s[k][p] != -1.0)
                    throw InvalidArgument("validate_context: " + grp + " s" + std::to_string(k) +
                                          " must be +-1 per sample");
        if (g.weight)
            for (long long p = 0; p < g.total_samples; ++p)
                if (!std::isfinite(g.weight[p]))
                    throw InvalidArgument("validate_context: " + grp + " weight must be finite");
        if (g.curv_k < 0) throw InvalidArgument("validate_context: " + grp + " negative curv_k");
        if (g.curv_k > 0) {
            if (dim != 2) throw InvalidArgument("validate_context: curvature windows are 2D only");
            if (g.curv_k > 64) throw InvalidArgument("validate_context: more than 64 curvature windows per DOF");
            for (long long q = 0; q < g.ndof * g.curv_k; ++q) {
                if (g.curv_slot[q] < -1 || g.curv_slot[q] > 3)
                    throw InvalidArgument("validate_context: " + grp + " curv_slot outside [-1, 3]");
                if (g.curv_slot[q] >= 0) {
                    if (!std::isfinite(g.curv_weight[q]))
                        throw InvalidArgument("validate_context: " + grp + " curv_weight must be finite");
                    check_index_range(g.curv_nodes + 4 * q, 4, N, grp, "curv_nodes");
                }
            }
        }
        if (g.dir_k < 0) throw InvalidArgument("validate_context: " + grp + " negative dir_k");
        if (g.dir_k > 0) {
            if (g.dir_k > 64) throw InvalidArgument("validate_context: more than 64 directional samples per DOF");
            validate_dir(dim, g.dir_nodes, g.dir_slot, g.dir_kind, g.dir_ref, g.dir_weight, g.dir_eps,
                         (long long)g.ndof * g.dir_k, N, grp);
        }
        if (g.entity_tag)
            for (long long d = 0; d < g.ndof; ++d) {
                const int t = g.entity_tag[d];
                if (t == 0) continue;
                if (t < 0 || t > kEntMaxTag)
                    throw InvalidArgument("validate_context: " + grp + " entity tag " + std::to_string(t) +
                                          " is not an entity kind");
                if (!g.entity_off || !g.entity_data)
                    throw InvalidArgument("validate_context: " + grp + " constrained DOFs need entity records");
                const long long off = g.entity_off[d];
                if (off < 0 || off >= g.entity_len)
                    throw InvalidArgument("validate_context: " + grp + " entity_off outside entity_data");
                if ((int)g.entity_data[off] != t)
                    throw InvalidArgument("validate_context: " + grp + " entity record's tag differs from entity_tag");
                entity_blob_check(g.entity_data + off, g.entity_len - off, dim);
            }
    }
    if (es->num_samples < 0) throw InvalidArgument("validate_context: energy_stencil negative size");
    if (es->w_rows != 1 && es->w_rows != es->num_samples)
        throw InvalidArgument("validate_context: energy_stencil W_inv size " + std::to_string(es->w_rows * dd) +
                              " != " + std::to_string(es->num_samples * dd));
    check_index_range(es->gc, es->num_samples, N, "energy_stencil", "gc");
    for (int k = 0; k < dim; ++k) {
        const std::string nm = "gn" + std::to_string(k);
        check_index_range(es->gn[k], es->num_samples, N, "energy_stencil", nm.c_str());
    }
    if (es->weight)
        for (long long i = 0; i < es->num_samples; ++i)
            if (!std::isfinite(es->weight[i]))
                throw InvalidArgument("validate_context: energy_stencil weight must be finite");
    if (es->dir_num < 0) throw InvalidArgument("validate_context: energy_stencil negative dir_num");
    if (es->dir_num > 0) {
        if (es->dir_num >= (1ll << 31)) throw InvalidArgument("validate_context: more than 2^31 directional samples");
        validate_dir(dim, es->dir_nodes, nullptr, es->dir_kind, es->dir_ref, es->dir_weight, es->dir_eps, es->dir_num, N,
                     "energy_stencil");
    }
    int gk = -1;  // the groups' directional eps must agree (one per session)
    for (int gi = 0; gi < n_groups; ++gi)
        if (groups[gi].dir_k > 0) {
            if (gk >= 0 && groups[gi].dir_eps != groups[gk].dir_eps)
                throw InvalidArgument("validate_context: the groups' dir_eps differ");
            gk = gi;
        }
}

// A directional table (reference src/directional.hpp:37-48): the used node
// slots of a sample of known kind in range, the DOF's slot (per-DOF tables)
// in [-1, 3] and below the kind's arity, finite reference vector, weight and
// eps. Samples of other kinds add nothing and are not checked further.
void validate_dir(int dim, const int32_t* nodes, const int32_t* slot, const int32_t* kind, const double* ref,
                  const double* wt, double eps, long long n, long long N, const std::string& grp)
{
    if (!std::isfinite(eps)) throw InvalidArgument("validate_context: " + grp + " dir_eps must be finite");
    for (long long q = 0; q < n; ++q) {
        const int sl = slot ? slot[q] : 0;
        if (sl < -1 || sl > 3) throw InvalidArgument("validate_context: " + grp + " dir_slot outside [-1, 3]");
        const int ar = dir_arity(kind[q]);
        if (sl < 0 || ar == 0) continue;
        if (sl >= ar)
            throw InvalidArgument("validate_context: " + grp + " dir_slot beyond its kind's " + std::to_string(ar) +
                                  " nodes");
        check_index_range(nodes + 4 * q, ar, N, grp, "dir_nodes");
        bool ok = std::isfinite(wt[q]);
        for (int k = 0; k < dim; ++k) ok = ok && std::isfinite(ref[(long long)dim * q + k]);
        if (!ok) throw InvalidArgument("validate_context: " + grp + " dir_ref / dir_weight must be finite");
    }
}

// A weight array that is absent or 1 everywhere leaves the unweighted sums.
bool unit_weights(const double* w, long long n)
{
    if (!w) return true;
    for (long long i = 0; i < n; ++i)
        if (w[i] != 1.0) return false;
    return true;
}

}  // namespace

// ------------------------------------------------------------------ session

void Session::ensure_timing_events(size_t n)
{
    while (tev.size() < n) {
        hipEvent_t e;
        EGGMI_HIP(hipEventCreate(&e));
        tev.push_back(e);
    }
}

Session::~Session()
{
    release_graphs();
    for (auto e : tev) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
}

void Session::create(int d, const eggmi_group* groups, int n_groups, const eggmi_stencil* es,
                     const eggmi_structured* st, const double* X, long long num_nodes, int device)
{
    if (d != 2 && d != 3) throw InvalidArgument("dim must be 2 or 3; got " + std::to_string(d));
    if (device < 0 || device > 3) throw InvalidArgument("device must be 0 (auto), 1 (gpu), 2 (cpu) or 3 (tables only)");
    dim = d;
    M = num_nodes;
    validate(d, groups, n_groups, es, num_nodes);

    // ---- structured remap (reference src/bindings.cpp:758-1101) ----
    const int nb = st->num_blocks;
    std::vector<std::array<int, 3>> shapes(nb, {1, 1, 1});
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < d; ++k) {
            shapes[b][k] = st->interior_shapes[b * d + k];
            if (shapes[b][k] < 1) throw InvalidArgument("structured: interior shape entries must be >= 1");
        }
    Layout& lay = fine.lay;
    lay.build(d, shapes);
    bgd.assign(nb, {});
    for (int b = 0; b < nb; ++b) {
        long long n = 1;
        for (int k = 0; k < d; ++k) n *= shapes[b][k];
        bgd[b].assign(st->block_global_dof[b], st->block_global_dof[b] + n);
        for (int v : bgd[b])
            if (v < 0 || v >= M)
                throw InvalidArgument("structured: block_global_dof entry " + std::to_string(v) + " out of range");
    }
    auto unravel = [&](int b, long long f, int* logical) {
        for (int k = d - 1; k >= 0; --k) {
            logical[k] = (int)(f % shapes[b][k]);
            f /= shapes[b][k];
        }
    };
    auto ravel = [&](int b, const int* logical) {
        long long f = 0;
        for (int k = 0; k < d; ++k) f = f * shapes[b][k] + logical[k];
        return f;
    };

    // pass 1: interior membership (gid -> (block, slot) list, CSR), owners,
    // initial scatter of X into every copy
    std::vector<int> owner_block(M, -1);
    std::vector<long long> owner_slot(M, -1);
    std::vector<int> mcount(M + 1, 0);
    for (int b = 0; b < nb; ++b)
        for (int g : bgd[b]) ++mcount[g + 1];
    for (long long g = 0; g < M; ++g) mcount[g + 1] += mcount[g];
    std::vector<int> mfill(mcount.begin(), mcount.end() - 1);
    std::vector<int> mblock(mcount[M]);
    std::vector<long long> mslot(mcount[M]);
    const long long nslots_blocks = lay.padded_nodes;
    std::vector<double> Xp((size_t)nslots_blocks * d, 0.0);
    for (int b = 0; b < nb; ++b) {
        int lg[3];
        for (long long f = 0; f < (long long)bgd[b].size(); ++f) {
            unravel(b, f, lg);
            const int gid = bgd[b][f];
            const long long sidx = lay.interior_slot(b, lg);
            mblock[mfill[gid]] = b;
            mslot[mfill[gid]++] = sidx;
            if (owner_block[gid] < 0) {
                owner_block[gid] = b;
                owner_slot[gid] = sidx;
            }
            for (int k = 0; k < d; ++k) Xp[(size_t)sidx * d + k] = X[(size_t)gid * d + k];
        }
    }
    // singular host override
    for (long long s = 0; s < st->n_sing; ++s) {
        const int ob = st->sing_block[s];
        if (ob < 0 || ob >= nb) throw InvalidArgument("structured: sing_block out of range");
        int lg[3];
        for (int k = 0; k < d; ++k) {
            lg[k] = st->sing_logical[s * d + k];
            if (lg[k] < 0 || lg[k] >= shapes[ob][k]) throw InvalidArgument("structured: sing_logical out of range");
        }
        const int gid = bgd[ob][ravel(ob, lg)];
        owner_block[gid] = ob;
        owner_slot[gid] = lay.interior_slot(ob, lg);
    }
    // pass 2: owner -> copy broadcasts
    std::vector<int> share_src, share_dst;
    for (int b = 0; b < nb; ++b) {
        int lg[3];
        for (long long f = 0; f < (long long)bgd[b].size(); ++f) {
            const int gid = bgd[b][f];
            if (owner_block[gid] == b) continue;
            unravel(b, f, lg);
            share_src.push_back((int)owner_slot[gid]);
            share_dst.push_back((int)lay.interior_slot(b, lg));
        }
    }
    // halo entries: ghost maps per block (these override interior entries)
    std::vector<std::unordered_map<int, long long>> ghost_map(nb);
    std::vector<int> halo_src, halo_dst;
    std::vector<std::vector<char>> used_ghost(nb);
    for (int b = 0; b < nb; ++b) used_ghost[b].assign(lay.block_padded_count(b), 0);
    for (long long e = 0; e < st->n_halo; ++e) {
        const int sb = st->halo_src_block[e], db = st->halo_dst_block[e];
        if (sb < 0 || sb >= nb || db < 0 || db >= nb) throw InvalidArgument("structured: halo block out of range");
        int sp[3], dp[3], sl[3];
        for (int k = 0; k < d; ++k) {
            sp[k] = st->halo_src_padded[e * d + k];
            dp[k] = st->halo_dst_padded[e * d + k];
            sl[k] = sp[k] - 1;
            if (sl[k] < 0 || sl[k] >= shapes[sb][k]) throw InvalidArgument("structured: halo_src_padded not interior");
            if (dp[k] < 0 || dp[k] > shapes[db][k] + 1) throw InvalidArgument("structured: halo_dst_padded out of range");
        }
        const int gid = bgd[sb][ravel(sb, sl)];
        const long long ds = lay.padded_slot(db, dp);
        fine_halo.sb.push_back(sb);
        fine_halo.db.push_back(db);
        fine_halo.sp.push_back({sp[0], d > 1 ? sp[1] : 0, d > 2 ? sp[2] : 0});
        fine_halo.dp.push_back({dp[0], d > 1 ? dp[1] : 0, d > 2 ? dp[2] : 0});
        ghost_map[db][gid] = ds;
        used_ghost[db][ds - lay.base[db]] = 1;
        halo_src.push_back((int)lay.padded_slot(sb, sp));
        halo_dst.push_back((int)ds);
    }
    // block-local lookup: ghost entries first (they override), then interior copies
    auto find_in_block = [&](int b, int gid, long long& slot) -> bool {
        auto it = ghost_map[b].find(gid);
        if (it != ghost_map[b].end()) {
            slot = it->second;
            return true;
        }
        for (int m = mcount[gid]; m < mcount[gid + 1]; ++m)
            if (mblock[m] == b) {
                slot = mslot[m];
                return true;
            }
        return false;
    };
    g2s.assign(M, -1);
    for (long long g = 0; g < M; ++g) {
        const int ob = owner_block[g];
        if (ob < 0) continue;
        long long slot;
        if (!find_in_block(ob, (int)g, slot))
            throw InvalidArgument("structured: owner_block names a block that does not contain its global node");
        g2s[g] = (int)slot;
    }
    // spare ghost-ring slots per block, in padded row-major order
    std::vector<std::vector<long long>> free_ghosts(nb);
    std::vector<size_t> cursor(nb, 0);
    for (int b = 0; b < nb; ++b) {
        const long long tot = lay.block_padded_count(b);
        for (long long f = 0; f < tot; ++f) {
            int pd[3];
            long long rem = f;
            bool ring = false;
            for (int k = d - 1; k >= 0; --k) {
                pd[k] = (int)(rem % (shapes[b][k] + 2));
                rem /= shapes[b][k] + 2;
            }
            for (int k = 0; k < d; ++k) ring = ring || pd[k] == 0 || pd[k] == shapes[b][k] + 1;
            if (ring && !used_ghost[b][f]) free_ghosts[b].push_back(lay.base[b] + f);
        }
    }
    std::vector<int> fan_src, fan_dst;
    long long n_overflow = 0;
    auto lookup = [&](int b, int gid, const char* what) -> long long {
        long long slot;
        if (find_in_block(b, gid, slot)) return slot;
        const int src = g2s[gid];
        if (src < 0)
            throw InvalidArgument(std::string("structured: ") + what + " references global node " +
                                  std::to_string(gid) + " absent from every block's interior map");
        if (cursor[b] < free_ghosts[b].size()) {
            slot = free_ghosts[b][cursor[b]++];
        } else {
            slot = (long long)(Xp.size() / d);
            for (int k = 0; k < d; ++k) Xp.push_back(Xp[(size_t)src * d + k]);
            ++n_overflow;
        }
        ghost_map[b][gid] = slot;
        fan_src.push_back(src);
        fan_dst.push_back((int)slot);
        return slot;
    };

    // DOF tables: synthesized interior DOFs get a mask bit; stored ones their
    // remapped occurrences (relative to the owner block)
    const int dd = d * d;
    std::vector<int> synth_slots;
    std::vector<int> gn_rows;  // [total][dim] stored neighbour slots
    Level& L = fine;
    L.st_off.assign(1, 0);
    bool w_uniform_all = true;
    std::vector<double> wu_ref;
    for (int gi = 0; gi < n_groups; ++gi)
        if (groups[gi].w_rows != 1) w_uniform_all = false;
    int kmax_curv = 0;  // curvature windows per special DOF (the widest group's)
    for (int gi = 0; gi < n_groups; ++gi) kmax_curv = std::max(kmax_curv, (int)groups[gi].curv_k);
    L.curv_k = kmax_curv;
    int kmax_dir = 0;  // directional windows per special DOF (the widest group's)
    for (int gi = 0; gi < n_groups; ++gi)
        if (groups[gi].dir_k > 0) {
            kmax_dir = std::max(kmax_dir, (int)groups[gi].dir_k);
            L.dir_eps = (float)groups[gi].dir_eps;
        }
    L.dir_k = kmax_dir;
    bool wt_unit_all = true;  // per-occurrence energy weights all 1 (or absent)
    for (int gi = 0; gi < n_groups; ++gi)
        if (!unit_weights(groups[gi].weight, groups[gi].total_samples)) wt_unit_all = false;
    if (w_uniform_all && n_groups > 0) {
        // all groups uniform: they must agree to share one uniform row
        for (int gi = 0; gi < n_groups; ++gi) {
            std::vector<double> w(groups[gi].W_inv, groups[gi].W_inv + dd);
            if (wu_ref.empty())
                wu_ref = w;
            else if (w != wu_ref)
                w_uniform_all = false;
        }
    }
    // entity records in FP32, one copy per distinct record (the DOFs of one
    // curve share it); returns its offset in the level's arena
    std::unordered_map<std::string, int> ent_seen;
    auto pack_entity = [&](const double* b, long long avail, int dim) -> int {
        const long long len = entity_blob_check(b, avail, dim);
        std::string key((const char*)b, (size_t)len * sizeof(double));
        auto it = ent_seen.find(key);
        if (it != ent_seen.end()) return it->second;
        if ((long long)L.con_blob.size() + len > (long long)std::numeric_limits<int>::max())
            throw InvalidArgument("structured: entity records exceed 2^31 values");
        const int off = (int)L.con_blob.size();
        for (long long i = 0; i < len; ++i) L.con_blob.push_back((float)b[i]);
        ent_seen.emplace(std::move(key), off);
        return off;
    };
    // constrained boundary DOFs (an entity to slide on) are stored last, so
    // the sweep's stored view excludes them and a kernel of their own takes them
    for (int pass = 0; pass < 2; ++pass)
    for (int gi = 0; gi < n_groups; ++gi) {
        const eggmi_group& g = groups[gi];
        long long off = 0;
        for (long long di = 0; di < g.ndof; ++di) {
            const bool con = g.entity_tag && g.entity_tag[di] != 0;
            bool curv = false;
            for (int j = 0; j < g.curv_k; ++j) curv = curv || g.curv_slot[di * g.curv_k + j] >= 0;
            for (int j = 0; j < g.dir_k; ++j) curv = curv || g.dir_slot[di * g.dir_k + j] >= 0;
            const bool special = con || curv;  // constrained, or with curvature / directional windows
            if (special != (pass == 1)) {
                off += g.P_of[di];
                continue;
            }
            const int gid = g.dof_idx[di];
            const int ob = owner_block[gid];
            if (ob < 0) throw InvalidArgument("structured: dof_idx " + std::to_string(gid) + " is in no block");
            const long long P = g.P_of[di];
            std::vector<long long> occ_slots;
            for (long long p = off; p < off + P; ++p) {
                occ_slots.push_back(lookup(ob, g.gc[p], "gc"));
                for (int k = 0; k < d; ++k) occ_slots.push_back(lookup(ob, g.gn[k][p], "gn"));
            }
            const long long dslot = lookup(ob, gid, "dof_idx");
            const int ib = g.interior_block[di];
            int lg[3] = {0, 0, 0};
            if (ib >= 0) {
                // a synthesized interior DOF, whichever path it takes below: it
                // ships no samples, names its own block slot strictly inside the
                // block, and its patch uses the uniform target with unit weights
                if (P != 0) throw InvalidArgument("structured: interior DOF with stored samples");
                if (ib >= nb) throw InvalidArgument("structured: interior_block out of range");
                for (int k = 0; k < d; ++k) {
                    lg[k] = g.interior_logical[di * d + k];
                    if (lg[k] < 1 || lg[k] > shapes[ib][k] - 2)
                        throw InvalidArgument("structured: interior_logical outside [1, n-2]");
                }
                if (lay.interior_slot(ib, lg) != dslot)
                    throw InvalidArgument("structured: interior DOF does not live at its block slot");
                if (!wt_unit_all)
                    throw InvalidArgument("structured: a synthesized interior DOF needs unit sample weights "
                                          "(ship its stored stencil instead)");
            }
            if (special) {
                if (con && ib >= 0)
                    throw InvalidArgument("structured: a synthesized interior DOF cannot be constrained");
                L.con_tag.push_back(con ? (uint8_t)g.entity_tag[di] : (uint8_t)0);
                L.con_off.push_back(con ? pack_entity(g.entity_data + g.entity_off[di],
                                                      g.entity_len - g.entity_off[di], d)
                                        : 0);
                for (int j = 0; j < kmax_curv; ++j) {
                    const bool live = j < g.curv_k && g.curv_slot[di * g.curv_k + j] >= 0;
                    L.curv_slot.push_back(live ? (int8_t)g.curv_slot[di * g.curv_k + j] : (int8_t)-1);
                    L.curv_w.push_back(live ? (float)g.curv_weight[di * g.curv_k + j] : 0.0f);
                    for (int m = 0; m < 4; ++m) {
                        const int nid = live ? g.curv_nodes[(di * g.curv_k + j) * 4 + m] : 0;
                        if (live && g2s[nid] < 0)
                            throw InvalidArgument("structured: a curvature window node is in no block");
                        L.curv_nodes.push_back(live ? g2s[nid] : 0);
                    }
                }
                for (int j = 0; j < kmax_dir; ++j) {
                    const long long q = di * g.dir_k + j;
                    const int ar = j < g.dir_k ? dir_arity(g.dir_kind[q]) : 0;
                    const bool live = ar > 0 && g.dir_slot[q] >= 0;
                    L.dir_slot.push_back(live ? (int8_t)g.dir_slot[q] : (int8_t)-1);
                    L.dir_kind.push_back(live ? (int8_t)g.dir_kind[q] : (int8_t)-1);
                    L.dir_w.push_back(live ? (float)g.dir_weight[q] : 0.0f);
                    for (int k = 0; k < d; ++k) L.dir_ref.push_back(live ? (float)g.dir_ref[(long long)d * q + k] : 0.0f);
                    for (int m = 0; m < 4; ++m) {
                        const int nid = live && m < ar ? g.dir_nodes[4 * q + m] : -1;
                        if (nid >= 0 && g2s[nid] < 0)
                            throw InvalidArgument("structured: a directional-sample node is in no block");
                        L.dir_nodes.push_back(nid >= 0 ? g2s[nid] : 0);
                    }
                }
                if (ib >= 0) {
                    // a synthesized interior DOF with curvature or directional windows takes the
                    // stored path: its patch expanded here in the reference order
                    // (cell bits outer, corner bits inner, axis 0 most significant;
                    // src/structured_patch.hpp:89-130) with the synthesized patch's
                    // identity target (src/structured_patch.hpp:175-213)
                    L.st_dof.push_back((int)dslot);
                    const int nc = 1 << d;
                    auto is_dof = [&](const int* c) {
                        for (int k = 0; k < d; ++k)
                            if (c[k] != lg[k]) return false;
                        return true;
                    };
                    for (int occ = 0; occ < nc * nc; ++occ) {
                        const int cell = occ / nc, o = occ % nc;
                        int corner[3] = {0, 0, 0}, ob2[3] = {0, 0, 0};
                        for (int k = 0; k < d; ++k) {
                            ob2[k] = (o >> (d - 1 - k)) & 1;
                            corner[k] = lg[k] - 1 + ((cell >> (d - 1 - k)) & 1) + ob2[k];
                        }
                        int role = is_dof(corner) ? 0 : -1;
                        L.st_gc.push_back((int)lay.interior_slot(ib, corner));
                        uint8_t meta = 0;
                        for (int k = 0; k < d; ++k) {
                            int nbr[3] = {corner[0], corner[1], corner[2]};
                            nbr[k] += ob2[k] == 0 ? 1 : -1;
                            gn_rows.push_back((int)lay.interior_slot(ib, nbr));
                            if (ob2[k]) meta |= (uint8_t)(8 << k);
                            if (role < 0 && is_dof(nbr)) role = k + 1;
                        }
                        L.st_meta.push_back((uint8_t)(meta | (uint8_t)(role + 1)));
                        if (!w_uniform_all)
                            for (int i = 0; i < dd; ++i) L.st_W.push_back(i % (d + 1) == 0 ? 1.0f : 0.0f);
                    }
                    L.st_off.push_back((int)L.st_gc.size());
                    off += P;
                    continue;
                }
            }
            if (ib >= 0) {
                synth_slots.push_back((int)dslot);
            } else {
                L.st_dof.push_back((int)dslot);
                for (long long p = off, q = 0; p < off + P; ++p, ++q) {
                    L.st_gc.push_back((int)occ_slots[q * (d + 1)]);
                    for (int k = 0; k < d; ++k) gn_rows.push_back((int)occ_slots[q * (d + 1) + 1 + k]);
                    if (g.role[p] < -1 || g.role[p] > d)
                        throw InvalidArgument("structured: role " + std::to_string(g.role[p]) + " outside [-1, " +
                                              std::to_string(d) + "]");
                    uint8_t meta = (uint8_t)(g.role[p] + 1);  // 0..D: 3 bits
                    for (int k = 0; k < d; ++k)
                        if (g.s[k][p] < 0) meta |= (uint8_t)(8 << k);
                    L.st_meta.push_back(meta);
                    if (!w_uniform_all) {
                        const double* w = g.w_rows == 1 ? g.W_inv : g.W_inv + p * dd;
                        for (int i = 0; i < dd; ++i) L.st_W.push_back((float)w[i]);
                    }
                    if (!wt_unit_all) L.st_wt.push_back(g.weight ? (float)g.weight[p] : 1.0f);
                }
                L.st_off.push_back((int)L.st_gc.size());
            }
            off += P;
        }
    }
    // gn stored [dim][total]
    {
        const long long total = (long long)L.st_gc.size();
        L.st_gn.assign((size_t)total * d, 0);
        for (long long q = 0; q < total; ++q)
            for (int k = 0; k < d; ++k) L.st_gn[(size_t)k * total + q] = gn_rows[(size_t)q * d + k];
    }
    if (w_uniform_all && !wu_ref.empty())
        for (int i = 0; i < dd; ++i) L.Wu[i] = (float)wu_ref[i];
    weighted = !wt_unit_all || !unit_weights(es->weight, es->num_samples);
    lay.overflow = n_overflow;
    const long long nslots = lay.n_slots();
    if ((long long)Xp.size() != nslots * d) throw Internal("structured: packed X size disagrees with the layout");
    if (nslots >= (1ll << 31)) throw InvalidArgument("structured: more than 2^31 node slots");
    L.mask.assign(nslots, 0);
    for (int s : synth_slots) L.mask[s] = 1;
    fine_free = L.mask;
    for (int s : L.st_dof) fine_free[s] = 1;
    // constrained DOFs stay out of the multigrid (its corrections would move
    // them off their entity): neither free for the hierarchy nor corrected
    for (size_t i = 0; i < L.con_tag.size(); ++i)
        if (L.con_tag[i]) fine_free[L.st_dof[L.st_dof.size() - L.con_tag.size() + i]] = 0;
    fine_halo.share_src.assign(share_src.begin(), share_src.end());
    fine_halo.share_dst.assign(share_dst.begin(), share_dst.end());
    L.n_synth = (long long)synth_slots.size();
    // copy table: halo | fan | share
    L.copy_src = halo_src;
    L.copy_dst = halo_dst;
    L.copy_src.insert(L.copy_src.end(), fan_src.begin(), fan_src.end());
    L.copy_dst.insert(L.copy_dst.end(), fan_dst.begin(), fan_dst.end());
    L.copy_src.insert(L.copy_src.end(), share_src.begin(), share_src.end());
    L.copy_dst.insert(L.copy_dst.end(), share_dst.begin(), share_dst.end());
    L.n_halo = (long long)halo_src.size();
    L.n_fan = (long long)fan_src.size();
    L.n_share = (long long)share_src.size();
    L.build_push();

    // ---- energy stencil: structured cells when the wire is the plain cell stencil ----
    std::vector<int> es_gc_h, es_gn_h;  // the explicit sample table otherwise (uploaded below)
    std::vector<float> es_s_h, es_W_h, es_wt_h;
    {
        const bool es_weighted = !unit_weights(es->weight, es->num_samples);
        bool canon = es->w_rows == 1 && !es_weighted;
        if (canon)
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j)
                    if (es->W_inv[i * d + j] != (i == j ? 1.0 : 0.0)) canon = false;
        const int nc = 1 << d;
        long long q = 0;
        for (int b = 0; b < nb && canon; ++b) {
            bool ok = true;
            for (int k = 0; k < d; ++k) ok = ok && shapes[b][k] >= 2;
            if (!ok) continue;
            int cdim[3] = {1, 1, 1};
            for (int k = 0; k < d; ++k) cdim[k] = shapes[b][k] - 1;
            const long long ncell = (long long)cdim[0] * cdim[1] * (d == 3 ? cdim[2] : 1);
            for (long long c = 0; c < ncell && canon; ++c) {
                int cl[3] = {0, 0, 0};
                long long rem = c;
                for (int k = d - 1; k >= 0; --k) {
                    cl[k] = (int)(rem % cdim[k]);
                    rem /= cdim[k];
                }
                for (int bb = 0; bb < nc && canon; ++bb, ++q) {
                    if (q >= es->num_samples) {
                        canon = false;
                        break;
                    }
                    int corner[3];
                    for (int k = 0; k < d; ++k) corner[k] = cl[k] + ((bb >> (d - 1 - k)) & 1);
                    if (es->gc[q] != bgd[b][ravel(b, corner)]) canon = false;
                    for (int k = 0; k < d && canon; ++k) {
                        int nbr[3] = {corner[0], corner[1], corner[2]};
                        const int bit = (bb >> (d - 1 - k)) & 1;
                        nbr[k] = cl[k] + (1 - bit);
                        if (es->gn[k][q] != bgd[b][ravel(b, nbr)]) canon = false;
                        if (es->s[k][q] != (bit ? -1.0 : 1.0)) canon = false;
                    }
                }
            }
        }
        if (q != es->num_samples) canon = false;
        L.structured_energy = canon;
        if (!canon) {
            const long long ns = es->num_samples;
            std::vector<int>& gc = es_gc_h;
            std::vector<int>& gn = es_gn_h;
            std::vector<float>& sv = es_s_h;
            std::vector<float>& W = es_W_h;
            gc.resize(ns);
            gn.resize((size_t)ns * d);
            sv.resize((size_t)ns * d);
            for (long long i = 0; i < ns; ++i) {
                gc[i] = g2s[es->gc[i]];
                for (int k = 0; k < d; ++k) {
                    gn[(size_t)k * ns + i] = g2s[es->gn[k][i]];
                    sv[(size_t)k * ns + i] = (float)es->s[k][i];
                }
                if (gc[i] < 0)
                    throw InvalidArgument("structured: an energy-stencil index references a global node absent "
                                          "from every block's interior map");
            }
            if (es->w_rows == 1) {
                for (int i = 0; i < dd; ++i) L.es_Wu[i] = es->W_inv[i];
            } else {
                W.assign(es->W_inv, es->W_inv + (size_t)ns * dd);
            }
            if (es_weighted)
                for (long long i = 0; i < ns; ++i) es_wt_h.push_back((float)es->weight[i]);
            L.es_n = ns;
        }
    }

    // ---- the energy reports' whole-sample directional table (owner slots) ----
    for (long long i = 0; i < es->dir_num; ++i) {
        const int ar = dir_arity(es->dir_kind[i]);
        L.es_dir_kind.push_back(ar ? (int8_t)es->dir_kind[i] : (int8_t)-1);
        L.es_dir_w.push_back(ar ? (float)es->dir_weight[i] : 0.0f);
        for (int k = 0; k < d; ++k) L.es_dir_ref.push_back(ar ? (float)es->dir_ref[(long long)d * i + k] : 0.0f);
        for (int m = 0; m < 4; ++m) {
            const int nid = m < ar ? es->dir_nodes[4 * i + m] : -1;
            if (nid >= 0 && g2s[nid] < 0)
                throw InvalidArgument("structured: a directional-sample node references a global node absent from "
                                      "every block's interior map");
            L.es_dir_nodes.push_back(nid >= 0 ? g2s[nid] : 0);
        }
    }

    // ---- device residency (host validation and remap above touch no device) ----
    if (device == 3) {  // tables only: index maps for inspection, no device state
        host_only = true;
        return;
    }
    if (device == 2)
        throw NoDevice("select_queue: no device available for device='cpu' (this library runs on the MI355X "
                       "GPU only). Use device='gpu' or 'auto'.");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw NoDevice("select_queue: no HIP device available for device='" +
                       std::string(device == 1 ? "gpu" : "auto") + "'.");
    EGGMI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    EGGMI_HIP(hipEventCreate(&ev0));
    EGGMI_HIP(hipEventCreate(&ev1));
    L.plane = (nslots + 63) / 64 * 64;
    L.upload_tables();
    L.make_tiles(tile_variant);
    if (!L.structured_energy) {
        L.d_es_gc.upload(es_gc_h);
        L.d_es_gn.upload(es_gn_h);
        L.d_es_s.upload(es_s_h);
        L.d_es_W.upload(es_W_h);
        L.d_es_wt.upload(es_wt_h);
    }
    if (es->dir_num > 0) {
        L.d_es_dir_nodes.upload(L.es_dir_nodes);
        L.d_es_dir_kind.upload(L.es_dir_kind);
        L.d_es_dir_ref.upload(L.es_dir_ref);
        L.d_es_dir_w.upload(L.es_dir_w);
        L.es_dir = DirView{(int)es->dir_num, (float)es->dir_eps, L.d_es_dir_nodes.p, nullptr, L.d_es_dir_kind.p,
                           L.d_es_dir_ref.p, L.d_es_dir_w.p};
    }
    std::vector<float> xs((size_t)L.plane * d, 0.0f);
    for (long long s = 0; s < nslots; ++s)
        for (int k = 0; k < d; ++k) xs[(size_t)k * L.plane + s] = (float)Xp[(size_t)s * d + k];
    L.x[0].upload(xs);
    L.x[1].alloc(xs.size());
    L.cur = 0;
    partials.alloc(4 * kReduceBlocks);  // reductions, then the directional table's blocks
}

void Session::halo(Level& L, float* x)
{
    launch_copy(x, L.plane, dim, L.d_copy_src.p, L.d_copy_dst.p, (long long)L.copy_src.size(), stream);
}

void Session::reduce(Level& L, const float* x, int kind, double delta, double* out_e, double* out_m, double* rep,
                     int cap, int* count)
{
    if (L.structured_energy) {
        CellReduceArgs a;
        a.x = x;
        a.plane = L.plane;
        a.blocks = L.blocks.p;
        a.cell_off = L.d_cell_off.p;
        a.nblocks = (int)L.lay.shape.size();
        a.ncells = L.ncells;
        a.delta = delta;
        launch_reduce_cells(dim, kind, a, partials.p, stream);
    } else {
        SampleReduceArgs a;
        a.x = x;
        a.plane = L.plane;
        a.ns = L.es_n;
        a.gc = L.d_es_gc.p;
        a.gn = L.d_es_gn.p;
        a.s = L.d_es_s.p;
        a.W = L.d_es_W.n ? L.d_es_W.p : nullptr;
        a.wt = L.d_es_wt.n ? L.d_es_wt.p : nullptr;
        for (int i = 0; i < 9; ++i) a.Wu[i] = L.es_Wu[i];
        a.delta = delta;
        launch_reduce_samples(dim, kind, a, partials.p, stream);
    }
    // whole-sample directional terms join the reported energy (reference src/sweep.hpp:1921)
    const int nd =
        launch_dir_reduce(dim, 0, L.es_dir, x, L.plane, nullptr, 0.0f, 0.0f, partials.p, kReduceBlocks, stream);
    launch_reduce_finish(partials.p, kReduceBlocks + nd, out_e, out_m, stream, rep, cap, count);
}

void Session::need_device() const
{
    if (host_only) throw InvalidArgument("session was built with device tables only (no device state)");
}

std::pair<std::vector<double>, std::vector<double>> Session::run_jacobi(int n_sweeps, int kind, double delta,
                                                                        double omega, int report_every)
{
    need_device();
    if (n_sweeps <= 0) return {{}, {}};
    const int k = report_every <= 0 ? n_sweeps : report_every;
    const int nrep = (n_sweeps + k - 1) / k;
    if ((int)d_reports.n < 2 * nrep) d_reports.alloc(2 * nrep);
    Level& L = fine;
    float* xin = L.x[L.cur].p;
    float* xout = L.x[1 - L.cur].p;
    const size_t bytes = (size_t)L.plane * dim * sizeof(float);
    EGGMI_HIP(hipEventRecord(ev0, stream));
    halo(L, xin);
    L.next_run(stream);
    // the write buffer must hold every slot no sweep writes (fixed nodes);
    // after a sweep it does, and it keeps the previous state of every DOF
    // (the learned proposal's previous-displacement input)
    if (!pair_valid) EGGMI_HIP(hipMemcpyAsync(xout, xin, bytes, hipMemcpyDeviceToDevice, stream));
    pair_valid = true;
    const bool propose = proposal != nullptr && kind != 1;
    const int mk = multi_k(L, propose);  // sweeps per multi-sweep launch (0: one per launch)
    int rep = 0;
    if (kernel_timing) ensure_timing_events(2 * n_sweeps);
    counters_begin();
    // graphs carry no per-launch events or counters: plain replays only
    const bool graphs_on = graph_chunk > 0 && !kernel_timing && !counting;
    long long launches = 0;
    int s = 0;
    while (s < n_sweeps) {
        const int next_rep = std::min(n_sweeps, (s / k + 1) * k);
        // a captured chunk repeats the cadence from its first sweep
        if (graphs_on && next_rep - s >= graph_chunk && graph_chunk % proposal_every == 0) {
            EGGMI_HIP(hipGraphLaunch(sweep_graph(xin == L.x[0].p ? 0 : 1, kind, (float)omega, (float)delta, propose),
                                     stream));
            s += graph_chunk;
            const int nl = chunk_launches(mk);
            launches += nl;
            if (nl % 2) std::swap(xin, xout);  // an odd launch count ends in the other buffer
        } else if (mk) {
            const int kk = std::min(mk, next_rep - s);
            if (kernel_timing) EGGMI_HIP(hipEventRecord(tev[2 * launches], stream));
            launch_multi(L, xin, xout, kk, kind, (float)omega, (float)delta, nullptr);
            if (kernel_timing) EGGMI_HIP(hipEventRecord(tev[2 * launches + 1], stream));
            std::swap(xin, xout);
            s += kk;
            ++launches;
        } else {
            SweepArgs a = L.sweep_args(xin, xout, (float)omega, (float)delta, nullptr);
            a.counters = counters_for(L);
            const bool here = propose && proposes_at(s);
            if (here) {
                a.mlp = proposal->view();
                a.prop_mode = proposal_mode;
                a.prop_scale = proposal_scale;
            }
            if (kernel_timing) EGGMI_HIP(hipEventRecord(tev[2 * launches], stream));
            launch_sweep(dim, kind, here ? proposal_variant() : tile_variant, a, stream);
            if (kernel_timing) EGGMI_HIP(hipEventRecord(tev[2 * launches + 1], stream));
            std::swap(xin, xout);
            ++s;
            ++launches;
        }
        if (s == next_rep) {
            reduce(L, xin, kind, delta, d_reports.p + rep, d_reports.p + nrep + rep);
            ++rep;
        }
    }
    EGGMI_HIP(hipEventRecord(ev1, stream));
    L.cur = (xin == L.x[0].p) ? 0 : 1;
    std::vector<double> hr(2 * nrep);
    EGGMI_HIP(hipMemcpyAsync(hr.data(), d_reports.p, sizeof(double) * 2 * nrep, hipMemcpyDeviceToHost, stream));
    EGGMI_HIP(hipStreamSynchronize(stream));
    counters_end();
    float ms = 0;
    EGGMI_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    last_ms = ms;
    last_launches = launches;
    last_kernel_ms = 0.0;
    if (kernel_timing)
        for (long long q = 0; q < launches; ++q) {
            float k = 0;
            EGGMI_HIP(hipEventElapsedTime(&k, tev[2 * q], tev[2 * q + 1]));
            last_kernel_ms += k;
        }
    return {std::vector<double>(hr.begin(), hr.begin() + nrep), std::vector<double>(hr.begin() + nrep, hr.end())};
}

void Session::release_graphs()
{
    for (auto& g : graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = SweepGraph {};
    }
    for (auto& g : fas_graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = FasGraph {};
    }
}

hipGraphExec_t Session::sweep_graph(int parity, int kind, float omega, float delta, bool propose)
{
    SweepGraph& g = graphs[parity];
    const void* mlp = propose ? (const void*)proposal->view().fwd : nullptr;
    if (g.exec && g.kind == kind && g.variant == tile_variant && g.chunk == graph_chunk && g.omega == omega &&
        g.delta == delta && g.mlp == mlp &&
        (!propose || (g.mode == proposal_mode && g.scale == proposal_scale && g.every == proposal_every)))
        return g.exec;
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = SweepGraph {};
    Level& L = fine;
    float* b0 = L.x[parity].p;
    float* b1 = L.x[1 - parity].p;
    hipGraph_t graph = nullptr;
    const int mk = multi_k(L, propose);
    EGGMI_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int i = 0, sw = 0; mk && sw < graph_chunk; ++i) {  // multi-sweep launches
            const int kk = std::min(mk, graph_chunk - sw);
            launch_multi(L, i % 2 == 0 ? b0 : b1, i % 2 == 0 ? b1 : b0, kk, kind, omega, delta, nullptr);
            sw += kk;
        }
        for (int i = 0; !mk && i < graph_chunk; ++i) {
            SweepArgs a = L.sweep_args(i % 2 == 0 ? b0 : b1, i % 2 == 0 ? b1 : b0, omega, delta, nullptr);
            const bool here = propose && proposes_at(i);
            if (here) {
                a.mlp = proposal->view();
                a.prop_mode = proposal_mode;
                a.prop_scale = proposal_scale;
            }
            launch_sweep(dim, kind, here ? proposal_variant() : tile_variant, a, stream);
        }
    } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    EGGMI_HIP(hipStreamEndCapture(stream, &graph));
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    EGGMI_HIP(ie);
    g.kind = kind;
    g.variant = tile_variant;
    g.chunk = graph_chunk;
    g.omega = omega;
    g.delta = delta;
    g.mlp = mlp;
    g.mode = proposal_mode;
    g.scale = proposal_scale;
    g.every = proposal_every;
    return g.exec;
}

void Level::ensure_synth_list()
{
    if (!syn_slot.empty() || n_synth == 0) return;
    for (size_t b = 0; b < lay.shape.size(); ++b) {
        const long long b0 = lay.base[b], b1 = b0 + lay.block_padded_count((int)b);
        for (long long sl = b0; sl < b1; ++sl)
            if (mask[sl]) {
                syn_slot.push_back((int)sl);
                syn_blk.push_back((int)b);
            }
    }
    d_syn_slot.upload(syn_slot);
    d_syn_blk.upload(syn_blk);
}

unsigned long long* Session::counters_for(const Level& L)
{
    if (!counting) return nullptr;
    constexpr long long per = (long long)kCounterSlots * kCounterStride;
    if (&L == &fine) return d_counters.p;
    for (size_t i = 0; i < coarse.size() && i + 1 < (size_t)kMaxCounterLevels; ++i)
        if (coarse[i].get() == &L) return d_counters.p + per * (i + 1);
    return nullptr;
}

void Session::counters_begin()
{
    if (!counting) return;
    const size_t n = (size_t)kCounterSlots * kCounterStride * kMaxCounterLevels;
    if (d_counters.n == 0) d_counters.alloc(n);
    EGGMI_HIP(hipMemsetAsync(d_counters.p, 0, n * sizeof(unsigned long long), stream));
}

// Called after the stream has been synchronised: sums the slot rows.
void Session::counters_end()
{
    if (!counting) {
        last_counters.clear();
        return;
    }
    const size_t per = (size_t)kCounterSlots * kCounterStride;
    std::vector<unsigned long long> h(per * kMaxCounterLevels);
    EGGMI_HIP(hipMemcpy(h.data(), d_counters.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    last_counters.assign((size_t)kCountersPerLevel * kMaxCounterLevels, 0);
    for (int l = 0; l < kMaxCounterLevels; ++l)
        for (int r = 0; r < kCounterSlots; ++r)
            for (int j = 0; j < 3; ++j)
                last_counters[(size_t)l * kCountersPerLevel + j] += h[l * per + (size_t)r * kCounterStride + j];
}

std::pair<double, double> Session::energy(int kind, double delta)
{
    need_device();
    if ((int)d_reports.n < 2) d_reports.alloc(2);
    reduce(fine, fine.x[fine.cur].p, kind, delta, d_reports.p, d_reports.p + 1);
    double h[2];
    EGGMI_HIP(hipMemcpyAsync(h, d_reports.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    EGGMI_HIP(hipStreamSynchronize(stream));
    return {h[0], h[1]};
}

void Session::set_tile_variant(int v)
{
    if (v < 0 || v >= num_tile_variants(dim)) throw InvalidArgument("tile variant out of range");
    if (proposal && tile_variant_proposal_partner(dim, v) < 0)
        throw InvalidArgument("tile variant " + std::to_string(v) + " has no learned-proposal kernel for its tile "
                              "table; unset the proposal first");
    tile_variant = v;
    if (host_only) return;
    release_graphs();  // they hold the tile table being replaced
    fine.make_tiles(v);
    for (size_t i = 0; i < coarse.size(); ++i)
        if (coarse[i]->x[0].n) apply_level_variant(i);
}

int Session::coarse_variant(size_t i) const
{
    return i < level_variant.size() && level_variant[i] >= 0 ? level_variant[i] : tile_variant;
}

void Session::apply_level_variant(size_t i)
{
    const int nv = num_tile_variants(dim), v = coarse_variant(i);
    coarse[i]->one_wg = v >= nv;
    coarse[i]->make_tiles(v % nv);
}

void Session::set_level_variants(const std::vector<int>& v)
{
    for (int x : v)
        if (x < -1 || x >= 2 * num_tile_variants(dim)) throw InvalidArgument("level tile variant out of range");
    level_variant = v;
    if (host_only) return;
    release_graphs();
    for (size_t i = 0; i < coarse.size(); ++i)
        if (coarse[i]->x[0].n) apply_level_variant(i);
}

std::vector<double> Session::autotune(int n_sweeps, double omega)
{
    need_device();
    if (n_sweeps < 1) throw InvalidArgument("autotune: n_sweeps must be >= 1");
    Level& L = fine;
    const size_t n = (size_t)L.plane * dim;
    DArr<float> save[2];
    for (int b = 0; b < 2; ++b) {
        save[b].alloc(n);
        EGGMI_HIP(hipMemcpyAsync(save[b].p, L.x[b].p, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    }
    const int cur0 = L.cur;
    const bool pair0 = pair_valid;
    const bool count0 = counting;
    const Mlp* prop0 = proposal;
    EGGMI_HIP(hipStreamSynchronize(stream));
    const Level::ConState con0 = L.save_con_state();
    counting = false;
    proposal = nullptr;
    const int nv = num_tile_variants(dim);
    std::vector<double> ms(nv, 0.0);
    const int msk0 = L.ms_sweeps;  // one-sweep kernels while their tiles are timed
    L.ms_sweeps = 0;
    try {
        for (int v = 0; v < nv; ++v) {
            if (prop0 && tile_variant_proposal_partner(dim, v) < 0) {
                ms[v] = std::numeric_limits<double>::infinity();  // not usable with the proposal set
                continue;
            }
            set_tile_variant(v);
            run_jacobi(1, 0, 0.0, omega, 1);  // warm the variant's code
            for (int b = 0; b < 2; ++b)
                EGGMI_HIP(hipMemcpyAsync(L.x[b].p, save[b].p, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
            L.cur = cur0;
            pair_valid = pair0;
            run_jacobi(n_sweeps, 0, 0.0, omega, n_sweeps);
            ms[v] = last_ms;
            for (int b = 0; b < 2; ++b)
                EGGMI_HIP(hipMemcpyAsync(L.x[b].p, save[b].p, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
            L.cur = cur0;
            pair_valid = pair0;
        }
    } catch (...) {
        counting = count0;
        proposal = prop0;
        L.ms_sweeps = msk0;
        throw;
    }
    EGGMI_HIP(hipStreamSynchronize(stream));
    L.restore_con_state(con0);
    int best = 0;
    for (int v = 1; v < nv; ++v)
        if (ms[v] < ms[best]) best = v;
    proposal = prop0;
    set_tile_variant(best);
    EGGMI_HIP(hipStreamSynchronize(stream));
    counting = count0;
    proposal = prop0;
    L.ms_sweeps = msk0;
    return ms;
}

std::vector<double> Session::autotune_multi(int n_sweeps, double omega, const std::vector<std::array<int, 4>>& cands,
                                            double base_ms)
{
    need_device();
    if (n_sweeps < 1) throw InvalidArgument("autotune_multi: n_sweeps must be >= 1");
    Level& L = fine;
    const size_t n = (size_t)L.plane * dim;
    DArr<float> save[2];
    for (int b = 0; b < 2; ++b) {
        save[b].alloc(n);
        EGGMI_HIP(hipMemcpyAsync(save[b].p, L.x[b].p, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    }
    const int cur0 = L.cur;
    const bool pair0 = pair_valid;
    const bool count0 = counting;
    EGGMI_HIP(hipStreamSynchronize(stream));
    counting = false;
    const std::array<int, 4> keep0 = {L.ms_sweeps, L.ms.ty, L.ms.tx, L.ms_kernel};
    std::vector<double> ms(cands.size(), std::numeric_limits<double>::infinity());
    auto restore = [&] {
        for (int b = 0; b < 2; ++b)
            EGGMI_HIP(hipMemcpyAsync(L.x[b].p, save[b].p, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        L.cur = cur0;
        pair_valid = pair0;
    };
    // each candidate is timed as one captured chunk of n_sweeps, replayed as
    // the timed runs replay them (host launch gaps would favour fewer launches)
    const int chunk0 = graph_chunk;
    graph_chunk = n_sweeps;
    try {
        for (size_t c = 0; c < cands.size(); ++c) {
            const auto& q = cands[c];
            try {
                set_multi_sweep(0, q[0], q[1], q[2], q[3]);
            } catch (const InvalidArgument&) {
                continue;  // not buildable (a closure outgrows the LDS): not timed
            }
            run_jacobi(n_sweeps, 0, 0.0, omega, n_sweeps);  // capture the chunk, warm the code
            restore();
            run_jacobi(n_sweeps, 0, 0.0, omega, n_sweeps);
            ms[c] = last_ms;
            restore();
        }
    } catch (...) {
        counting = count0;
        graph_chunk = chunk0;
        restore();
        set_multi_sweep(0, 0, 0, 0, 0);
        throw;
    }
    EGGMI_HIP(hipStreamSynchronize(stream));
    counting = count0;
    graph_chunk = chunk0;
    release_graphs();
    size_t best = 0;
    for (size_t c = 1; c < cands.size(); ++c)
        if (ms[c] < ms[best]) best = c;
    if (!cands.empty() && ms[best] < base_ms)
        set_multi_sweep(0, cands[best][0], cands[best][1], cands[best][2], cands[best][3]);
    else if (keep0[0] > 1)
        set_multi_sweep(0, keep0[0], keep0[1], keep0[2], keep0[3]);
    else
        set_multi_sweep(0, 0, 0, 0, 0);
    return ms;
}

void Session::get_X(double* out) const
{
    need_device();
    const Level& L = fine;
    std::vector<float> xs((size_t)L.plane * dim);
    EGGMI_HIP(hipMemcpy(xs.data(), L.x[L.cur].p, xs.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (long long g = 0; g < M; ++g) {
        const int s = g2s[g];
        for (int k = 0; k < dim; ++k) out[(size_t)g * dim + k] = s < 0 ? 0.0 : (double)xs[(size_t)k * L.plane + s];
    }
}

void Session::set_X(const double* X)
{
    need_device();
    Level& L = fine;
    std::vector<float> xs((size_t)L.plane * dim);
    EGGMI_HIP(hipMemcpy(xs.data(), L.x[L.cur].p, xs.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (long long g = 0; g < M; ++g) {
        const int s = g2s[g];
        if (s < 0) continue;
        for (int k = 0; k < dim; ++k) xs[(size_t)k * L.plane + s] = (float)X[(size_t)g * dim + k];
    }
    EGGMI_HIP(hipMemcpy(L.x[L.cur].p, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice));
    halo(L, L.x[L.cur].p);
    EGGMI_HIP(hipStreamSynchronize(stream));
    pair_valid = false;
    L.reset_seeds();
}

std::vector<float> Session::packed() const
{
    need_device();
    const Level& L = fine;
    const long long ns = L.lay.n_slots();
    std::vector<float> xs((size_t)L.plane * dim), out((size_t)ns * dim);
    EGGMI_HIP(hipMemcpy(xs.data(), L.x[L.cur].p, xs.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int k = 0; k < dim; ++k)
        std::memcpy(out.data() + (size_t)k * ns, xs.data() + (size_t)k * L.plane, ns * sizeof(float));
    return out;
}

}  // namespace eggmi
