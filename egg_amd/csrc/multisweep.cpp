// Host tables of the multi-sweep kernel (multisweep2d.hip): per core tile,
// the dependency closure of its k-th sweep, laid out for LDS staging.
//
// Dependencies. A synthesized DOF's update reads its 3x3 padded
// neighbourhood; a stored DOF's reads the corner and neighbour slots of its
// stored occurrences. A slot that is a replica (ghost, share or fan copy)
// holds its root's state after every sweep's push, so the dependency graph
// runs over roots. The closure of a core C is the set of roots within graph
// distance k of C; DOFs within distance k-1 are recomputed in LDS.
//
// Layout. For every block whose synthesized DOFs are recomputed, the
// bounding rectangle of those DOFs grown by one node is staged as a region
// (row addressing, as the one-sweep kernel's tile); closure roots outside
// every region are extra entries. Ghost slots inside a region whose root is
// staged too get that root's new state after each sub-sweep (push pairs).
// The DOFs to recompute are listed per tile, nearest first (synthesized ones
// with their LDS entry and region row width, stored ones with their
// occurrences mapped to LDS entries), so a sub-sweep's work is a dense prefix.
#include "session.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace eggmi
{

namespace
{
struct Geo {  // slot -> block and padded coordinates (2D)
    const Layout& lay;
    explicit Geo(const Layout& l) : lay(l) {}
    int block(long long slot) const
    {
        if (slot < 0 || slot >= lay.padded_nodes) return -1;
        const auto it = std::upper_bound(lay.base.begin(), lay.base.end(), slot);
        return (int)(it - lay.base.begin()) - 1;
    }
    void padded(int b, long long slot, int& p0, int& p1) const
    {
        const long long off = slot - lay.base[b];
        p0 = (int)(off / lay.stride[b][0]);
        p1 = (int)(off % lay.stride[b][0]);
    }
};
}  // namespace

void Level::build_multi(int k, int ty, int tx, bool upload)
{
    if (lay.dim != 2) throw InvalidArgument("multi-sweep: 2D levels only");
    if (k < 1 || k > kMsMaxK) throw InvalidArgument("multi-sweep: sweeps per launch must be in [1, 8]");
    if (ty < 1 || tx < 1 || ty > 256 || tx > 256) throw InvalidArgument("multi-sweep: tile sides must be in [1, 256]");
    if (!con_tag.empty())
        throw InvalidArgument("multi-sweep: special stored DOFs (constrained, curvature or directional windows) "
                              "take their own launch");
    if (!st_wt.empty()) throw InvalidArgument("multi-sweep: per-sample energy weights are not supported");
    const Geo geo(lay);
    const long long ns = lay.n_slots();
    const int nb = (int)lay.shape.size();
    const long long total_occ = (long long)st_gc.size();
    constexpr int G = stored_lanes(2);

    std::vector<int> root(ns);
    std::iota(root.begin(), root.end(), 0);
    if (!rp_off.empty())
        for (long long r = 0; r < ns; ++r)
            for (int q = rp_off[r]; q < rp_off[r + 1]; ++q) root[rp_dst[q]] = (int)r;
    std::vector<int> st_of(ns, -1);
    for (size_t d = 0; d < st_dof.size(); ++d) st_of[st_dof[d]] = (int)d;
    for (size_t d = 0; d < st_dof.size(); ++d)
        if (root[st_dof[d]] != st_dof[d]) throw Internal("multi-sweep: a stored DOF's slot is a replica");
    auto is_dof = [&](int s) { return mask[s] != 0 || st_of[s] >= 0; };
    // roots read by DOF c's update
    auto stencil = [&](int c, auto&& f) {
        if (mask[c]) {
            const long long rs = lay.stride[geo.block(c)][0];
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj) f(root[c + di * rs + dj]);
        } else {
            const int d = st_of[c];
            for (int p = st_off[d]; p < st_off[d + 1]; ++p) {
                f(root[st_gc[p]]);
                for (int a = 0; a < 2; ++a) f(root[st_gn[(size_t)a * total_occ + p]]);
            }
        }
    };

    // core tiles: ty x tx logical nodes of a block's synthesized range
    // [1, n-2] (the face rows 0 and n-1 join the tiles next to them); a
    // stored DOF joins the tile of its slot's block nearest to it (ghost
    // positions clamped inward; a slot outside every block goes with its
    // first stencil slot inside one)
    std::vector<int> tile_first(nb + 1, 0), tiles_i(nb), tiles_j(nb);
    for (int b = 0; b < nb; ++b) {
        tiles_i[b] = std::max(1, (lay.shape[b][0] - 2 + ty - 1) / ty);
        tiles_j[b] = std::max(1, (lay.shape[b][1] - 2 + tx - 1) / tx);
        tile_first[b + 1] = tile_first[b] + tiles_i[b] * tiles_j[b];
    }
    const int n_tiles_all = tile_first[nb];
    std::vector<std::vector<int>> core(n_tiles_all);
    auto tile_of = [&](int b, int l0, int l1) {
        const int t0 = std::clamp((l0 - 1) / ty, 0, tiles_i[b] - 1), t1 = std::clamp((l1 - 1) / tx, 0, tiles_j[b] - 1);
        return tile_first[b] + t0 * tiles_j[b] + t1;
    };
    auto tile_at = [&](int b, int p0, int p1) {  // padded coordinates
        return tile_of(b, std::clamp(p0 - 1, 0, lay.shape[b][0] - 1), std::clamp(p1 - 1, 0, lay.shape[b][1] - 1));
    };
    for (int b = 0; b < nb; ++b)
        for (int i = 0; i < lay.shape[b][0]; ++i)
            for (int j = 0; j < lay.shape[b][1]; ++j) {
                const int lg[3] = {i, j, 0};
                const long long s = lay.interior_slot(b, lg);
                if (mask[s]) core[tile_of(b, i, j)].push_back((int)s);
            }
    for (size_t d = 0; d < st_dof.size(); ++d) {
        long long s = st_dof[d];
        if (geo.block(s) < 0) {
            s = -1;
            for (int p = st_off[d]; p < st_off[d + 1] && s < 0; ++p)
                if (geo.block(st_gc[p]) >= 0) s = st_gc[p];
            if (s < 0) throw Internal("multi-sweep: a stored DOF with no slot inside a block");
        }
        const int b = geo.block(s);
        int p0, p1;
        geo.padded(b, s, p0, p1);
        core[tile_at(b, p0, p1)].push_back(st_dof[d]);
    }

    MsHost H;
    H.k = k;
    H.ty = ty;
    H.tx = tx;
    H.G = G;
    std::vector<int> dist(ns, -1), lds_of(ns, -1);
    std::vector<int> touched, front, next;
    struct StEntry {
        int dist, slot;
    };
    std::vector<std::vector<StEntry>> st_entries;  // per emitted tile
    const int max_stride = (int)(kMsMaxLdsBytes / (2 * 2 * sizeof(float)));
    int rounds = 1;
    for (int t = 0; t < n_tiles_all; ++t) {
        if (core[t].empty()) continue;
        // ---- breadth-first distances over roots
        touched.clear();
        front.clear();
        for (int c : core[t]) {
            dist[c] = 0;
            touched.push_back(c);
            front.push_back(c);
        }
        for (int r = 0; r < k; ++r) {
            next.clear();
            for (int c : front)
                stencil(c, [&](int q) {
                    if (dist[q] >= 0) return;
                    dist[q] = r + 1;
                    touched.push_back(q);
                    if (r + 1 <= k - 1 && is_dof(q)) next.push_back(q);
                });
            front.swap(next);
        }
        // ---- regions: per block, the bounding box of its recomputed
        // synthesized DOFs grown by one node (their 3x3 neighbourhoods)
        struct Box {
            int b, lo0, hi0, lo1, hi1;
        };
        std::vector<Box> boxes;
        for (int c : touched) {
            if (!mask[c] || dist[c] > k - 1) continue;
            const int b = geo.block(c);
            int p0, p1;
            geo.padded(b, c, p0, p1);
            auto it = std::find_if(boxes.begin(), boxes.end(), [&](const Box& e) { return e.b == b; });
            if (it == boxes.end()) {
                boxes.push_back({b, p0, p0, p1, p1});
            } else {
                it->lo0 = std::min(it->lo0, p0);
                it->hi0 = std::max(it->hi0, p0);
                it->lo1 = std::min(it->lo1, p1);
                it->hi1 = std::max(it->hi1, p1);
            }
        }
        std::sort(boxes.begin(), boxes.end(), [](const Box& x, const Box& y) { return x.b < y.b; });
        MsTile T {};
        T.reg0 = (int)H.regions.size();
        T.nreg = (int)boxes.size();
        int n_lds = 0;
        std::vector<int> staged;  // slots given an LDS entry (reset below)
        for (const Box& o : boxes) {
            const int b = o.b;
            const int r0 = std::max(0, o.lo0 - 1), r1 = std::min(lay.shape[b][0] + 1, o.hi0 + 1);
            const int q0 = std::max(0, o.lo1 - 1), q1 = std::min(lay.shape[b][1] + 1, o.hi1 + 1);
            MsRegion R {};
            const int pr[3] = {r0, q0, 0};
            R.base = lay.padded_slot(b, pr);
            R.st0 = lay.stride[b][0];
            R.h = r1 - r0 + 1;
            R.w = q1 - q0 + 1;
            R.lds = n_lds;
            for (int i = 0; i < R.h; ++i)
                for (int j = 0; j < R.w; ++j) {
                    const long long s = R.base + (long long)i * R.st0 + j;
                    lds_of[s] = n_lds + i * R.w + j;
                    staged.push_back((int)s);
                }
            n_lds += R.h * R.w;
            H.regions.push_back(R);
        }
        // ---- extra entries: closure roots outside every region
        T.ext0 = (int)H.ext.size();
        for (int c : touched)
            if (lds_of[c] < 0) {
                lds_of[c] = n_lds++;
                staged.push_back(c);
                H.ext.push_back(c);
            }
        T.next = (int)H.ext.size() - T.ext0;
        T.n_lds = n_lds;
        if (n_lds > max_stride)
            throw InvalidArgument("multi-sweep: a tile's closure (" + std::to_string(n_lds) +
                                  " nodes) outgrows the LDS; take smaller tiles or fewer sweeps per launch");
        // ---- synthesized entries within distance k-1, nearest first; each
        // with its LDS entry, its region's row width and its store slot
        {
            std::vector<std::pair<int, int>> sy;  // (distance, slot)
            for (int c : touched)
                if (mask[c] && dist[c] <= k - 1) sy.emplace_back(dist[c], c);
            std::sort(sy.begin(), sy.end());
            T.syn0 = (int)H.syn.size();
            for (int r = 0; r < kMsMaxK; ++r) {
                int n = 0;
                for (auto& e : sy) n += e.first <= r ? 1 : 0;
                T.syn_cnt[r] = n;
            }
            for (auto& e : sy) {
                const int c = e.second, li = lds_of[c];
                int w = -1;
                for (int ri = 0; ri < T.nreg; ++ri) {
                    const MsRegion& R = H.regions[T.reg0 + ri];
                    if (li >= R.lds && li < R.lds + R.h * R.w) w = R.w;
                }
                if (w < 0 || li > 0xFFFF || w > 0x7FFF) throw Internal("multi-sweep: a synthesized entry outside its region");
                int2 en;
                en.x = li | (w << 16);
                en.y = c;
                H.syn.push_back(en);
            }
        }
        // ---- stored entries within distance k-1, nearest first
        std::vector<StEntry> se;
        for (int c : touched)
            if (st_of[c] >= 0 && dist[c] <= k - 1) se.push_back({dist[c], c});
        std::sort(se.begin(), se.end(), [](const StEntry& x, const StEntry& y) {
            return x.dist != y.dist ? x.dist < y.dist : x.slot < y.slot;
        });
        for (int r = 0; r < kMsMaxK; ++r) {
            int n = 0;
            for (auto& e : se) n += e.dist <= r ? 1 : 0;
            T.st_cnt[r] = r < k ? n : (int)se.size();
        }
        for (auto& e : se) {
            const int d = st_of[e.slot];
            rounds = std::max(rounds, (st_off[d + 1] - st_off[d] + G - 1) / G);
            for (int p = st_off[d]; p < st_off[d + 1]; ++p) {
                if (lds_of[root[st_gc[p]]] < 0) throw Internal("multi-sweep: a stored stencil node is not staged");
                for (int a = 0; a < 2; ++a)
                    if (lds_of[root[st_gn[(size_t)a * total_occ + p]]] < 0)
                        throw Internal("multi-sweep: a stored stencil node is not staged");
            }
        }
        // ---- push pairs: region ghosts whose root DOF is staged
        T.push0 = (int)H.push_src.size();
        for (int ri = 0; ri < T.nreg; ++ri) {
            const MsRegion& R = H.regions[T.reg0 + ri];
            for (int i = 0; i < R.h; ++i)
                for (int j = 0; j < R.w; ++j) {
                    const long long s = R.base + (long long)i * R.st0 + j;
                    const int rt = root[s];
                    if (rt != s && is_dof(rt) && lds_of[rt] >= 0) {
                        H.push_src.push_back(lds_of[rt]);
                        H.push_dst.push_back(lds_of[s]);
                    }
                }
        }
        T.npush = (int)H.push_src.size() - T.push0;
        // ---- accounting: DOF updates over the k sub-sweeps
        H.core_dofs += (long long)core[t].size();
        for (int r = 0; r < k; ++r) H.updates += T.syn_cnt[r] + T.st_cnt[r];
        H.tiles.push_back(T);
        st_entries.push_back(std::move(se));
        for (int c : touched) dist[c] = -1;
        for (int s : staged) lds_of[s] = -1;
    }
    // ---- second pass (the lane layout's rounds are known now): the stored
    // occurrences in the padded lane layout, each node as its tile's LDS entry
    H.rounds = rounds;
    long long n_st = 0;
    for (auto& se : st_entries) n_st += (long long)se.size();
    H.n_st = n_st;
    H.total = n_st * rounds * G;
    H.st_slot.assign(n_st, 0);
    H.st_lds.assign(n_st, 0);
    H.gc.assign(H.total, 0);
    H.gn.assign(2 * H.total, 0);
    H.meta.assign(H.total, kStoredPad);
    if (!st_W.empty()) H.W.assign((size_t)4 * H.total, 0.0f);
    long long e = 0;
    for (size_t ti = 0; ti < H.tiles.size(); ++ti) {
        MsTile& T = H.tiles[ti];
        T.st0 = (int)e;
        // the tile's LDS map again
        std::vector<int> staged;
        for (int ri = 0; ri < T.nreg; ++ri) {
            const MsRegion& R = H.regions[T.reg0 + ri];
            for (int i = 0; i < R.h; ++i)
                for (int j = 0; j < R.w; ++j) {
                    const long long s = R.base + (long long)i * R.st0 + j;
                    lds_of[s] = R.lds + i * R.w + j;
                    staged.push_back((int)s);
                }
        }
        for (int x = 0; x < T.next; ++x) {
            const int s = H.ext[T.ext0 + x];
            lds_of[s] = T.n_lds - T.next + x;
            staged.push_back(s);
        }
        for (auto& en : st_entries[ti]) {
            const int d = st_of[en.slot];
            H.st_slot[e] = en.slot;
            H.st_lds[e] = lds_of[en.slot];
            for (int p = st_off[d]; p < st_off[d + 1]; ++p) {
                const int o = p - st_off[d];
                const long long q = (e * rounds + o / G) * G + o % G;
                H.gc[q] = lds_of[root[st_gc[p]]];
                for (int a = 0; a < 2; ++a) H.gn[(size_t)a * H.total + q] = lds_of[root[st_gn[(size_t)a * total_occ + p]]];
                H.meta[q] = st_meta[p];
                if (!H.W.empty())
                    for (int x = 0; x < 4; ++x) H.W[(size_t)4 * q + x] = st_W[(size_t)4 * p + x];
            }
            ++e;
        }
        for (int s : staged) lds_of[s] = -1;
    }
    // every index the kernel forms stays inside its tile's LDS and the store
    int stride = 0;
    const auto bad = [](const char* what) { throw Internal(std::string("multi-sweep tables: ") + what); };
    for (const MsTile& T : H.tiles) {
        stride = std::max(stride, T.n_lds);
        for (int ri = 0; ri < T.nreg; ++ri) {
            const MsRegion& R = H.regions[T.reg0 + ri];
            if (R.lds < 0 || R.lds + R.h * R.w > T.n_lds - T.next) bad("region outside the tile's LDS");
            if (R.base < 0 || R.base + (long long)(R.h - 1) * R.st0 + R.w > lay.padded_nodes) bad("region slots");
        }
        for (int i = 0; i < T.syn_cnt[k - 1]; ++i) {  // each entry's 3x3 inside its region
            const int2 en = H.syn[T.syn0 + i];
            const int li = en.x & 0xFFFF, w = en.x >> 16;
            bool ok = false;
            for (int ri = 0; ri < T.nreg && !ok; ++ri) {
                const MsRegion& R = H.regions[T.reg0 + ri];
                const int e = li - R.lds, row = e / R.w, col = e % R.w;
                ok = e >= 0 && e < R.h * R.w && R.w == w && row >= 1 && row <= R.h - 2 && col >= 1 && col <= R.w - 2 &&
                     R.base + (long long)row * R.st0 + col == en.y;
            }
            if (!ok || !mask[en.y]) bad("a synthesized entry's 3x3 leaves its region");
        }
        for (int p = 0; p < T.npush; ++p)
            if (std::min(H.push_src[T.push0 + p], H.push_dst[T.push0 + p]) < 0 ||
                std::max(H.push_src[T.push0 + p], H.push_dst[T.push0 + p]) >= T.n_lds)
                bad("push pair");
        for (int i = 0; i < T.st_cnt[k - 1]; ++i) {
            const long long d = T.st0 + i;
            if (H.st_lds[d] < 0 || H.st_lds[d] >= T.n_lds || H.st_slot[d] < 0 || H.st_slot[d] >= ns) bad("stored entry");
            for (long long q = d * rounds * G; q < (d + 1) * rounds * G; ++q) {
                if (H.meta[q] & kStoredPad) continue;
                if (std::min({H.gc[q], H.gn[q], H.gn[H.total + q]}) < 0 ||
                    std::max({H.gc[q], H.gn[q], H.gn[H.total + q]}) >= T.n_lds)
                    bad("stored occurrence");
            }
        }
        for (int x = 0; x < T.next; ++x)
            if (H.ext[T.ext0 + x] < 0 || H.ext[T.ext0 + x] >= ns) bad("extra node");
    }
    H.lds_stride = (stride + 3) & ~3;
    ms = std::move(H);
    if (upload) {
        d_ms_tiles.upload(ms.tiles);
        d_ms_regions.upload(ms.regions);
        d_ms_ext.upload(ms.ext);
        d_ms_syn.upload(ms.syn);
        d_ms_st_slot.upload(ms.st_slot);
        d_ms_st_lds.upload(ms.st_lds);
        d_ms_gc.upload(ms.gc);
        d_ms_gn.upload(ms.gn);
        d_ms_meta.upload(ms.meta);
        d_ms_W.upload(ms.W);
        d_ms_push_src.upload(ms.push_src);
        d_ms_push_dst.upload(ms.push_dst);
    }
}

MsArgs Level::multi_args(int k) const
{
    MsArgs m;
    m.k = k;
    m.n_tiles = (int)ms.tiles.size();
    m.lds_stride = ms.lds_stride;
    m.tiles = d_ms_tiles.p;
    m.regions = d_ms_regions.p;
    m.ext = d_ms_ext.p;
    m.syn = d_ms_syn.p;
    m.st_slot = d_ms_st_slot.p;
    m.st.n = (int)ms.n_st;
    m.st.rounds = ms.rounds;
    m.st.dof = d_ms_st_lds.p;
    m.st.total = ms.total;
    m.st.gc = d_ms_gc.p;
    m.st.gn = d_ms_gn.p;
    m.st.meta = d_ms_meta.p;
    m.st.W = ms.W.empty() ? nullptr : d_ms_W.p;
    m.st.wt = nullptr;
    m.push_src = d_ms_push_src.p;
    m.push_dst = d_ms_push_dst.p;
    return m;
}

// ------------------------------------------------------------------ session

Level& Session::level_ref(int level)
{
    if (level == 0) return fine;
    build_hierarchy();
    if (level < 0 || level > (int)coarse.size())
        throw InvalidArgument("level " + std::to_string(level) + " outside [0, " + std::to_string(coarse.size()) + "]");
    return *coarse[level - 1];
}

void Session::set_multi_sweep(int level, int k, int ty, int tx, int kernel)
{
    Level& L = level_ref(level);
    if (kernel < 0 || kernel >= num_ms_variants()) throw InvalidArgument("multi-sweep kernel variant out of range");
    if (k > 1 && level > 0 && !host_only && !L.x[0].n)
        throw InvalidArgument("multi-sweep: coarse level " + std::to_string(level) + " is not prepared (run_fas first)");
    if (!host_only) release_graphs();  // captured launches hold the tables being replaced
    if (k <= 1) {
        L.ms_sweeps = 0;
        return;
    }
    if (L.ms.k != k || L.ms.ty != ty || L.ms.tx != tx) {
        L.ms_sweeps = 0;  // stays off if the build throws
        L.build_multi(k, ty, tx, !host_only);
    }
    L.ms_kernel = kernel;
    L.ms_sweeps = k;
}

void Session::launch_multi(Level& L, const float* xin, float* xout, int k, int kind, float omega, float delta,
                           const float* tau)
{
    SweepArgs a = L.sweep_args(xin, xout, omega, delta, tau);
    a.counters = counters_for(L);
    launch_multi_sweep(kind, L.ms_kernel, a, L.multi_args(k), stream);
}

std::vector<double> Session::autotune_levels_multi(int n_sweeps, double omega, int max_levels, int kind,
                                                   const std::vector<std::array<int, 4>>& cands)
{
    need_device();
    if (n_sweeps < 1) throw InvalidArgument("autotune_levels_multi: n_sweeps must be >= 1");
    if (kind == 1) throw InvalidArgument("autotune_levels_multi: FAS smooths the barrier phases only");
    const size_t n_coarse = prepare_fas(max_levels);
    const size_t nc = cands.size() + 1;
    std::vector<double> ms(n_coarse * nc, -1.0);
    if (n_coarse == 0) return ms;
    const bool count0 = counting;
    const long long launches0 = last_launches;
    counting = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto drop = [&] {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    };
    try {
        Level& F = fine;
        const float* x = F.x[F.cur].p;
        halo(F, F.x[F.cur].p);
        grad(F, x, F.grad.p, nullptr, kind, false, true);
        restrict_to(F, x, *coarse[0], kind);
        for (size_t i = 0; i < n_coarse; ++i) {
            Level& L = *coarse[i];
            const long long n = (long long)L.plane * dim;
            auto reset = [&] {  // both Jacobi buffers <- the restricted state
                launch_copy2(L.x[0].p, L.x[1].p, L.rx.p, n, stream);
                L.cur = 0;
            };
            // one smoothing call, timed as the captured graph a cycle replays
            auto time_it = [&] {
                reset();
                EGGMI_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                try {
                    smooth(L, n_sweeps, kind, (float)omega, L.tau.p);
                } catch (...) {
                    (void)hipStreamEndCapture(stream, &graph);
                    throw;
                }
                EGGMI_HIP(hipStreamEndCapture(stream, &graph));
                EGGMI_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                EGGMI_HIP(hipGraphLaunch(exec, stream));  // warm the code
                reset();
                EGGMI_HIP(hipEventRecord(ev0, stream));
                EGGMI_HIP(hipGraphLaunch(exec, stream));
                EGGMI_HIP(hipEventRecord(ev1, stream));
                EGGMI_HIP(hipEventSynchronize(ev1));
                float t = 0;
                EGGMI_HIP(hipEventElapsedTime(&t, ev0, ev1));
                drop();
                reset();
                return (double)t;
            };
            L.ms_sweeps = 0;  // the level's one-sweep setting first
            ms[i * nc] = time_it();
            int best = -1;
            for (size_t c = 0; c < cands.size(); ++c) {
                const auto& q = cands[c];
                try {
                    set_multi_sweep((int)i + 1, q[0], q[1], q[2], q[3]);
                } catch (const InvalidArgument&) {
                    continue;
                }
                ms[i * nc + 1 + c] = time_it();
                if (ms[i * nc + 1 + c] < ms[i * nc + (best < 0 ? 0 : 1 + best)]) best = (int)c;
            }
            if (best >= 0)
                set_multi_sweep((int)i + 1, cands[best][0], cands[best][1], cands[best][2], cands[best][3]);
            else
                L.ms_sweeps = 0;
            if (i + 1 < n_coarse) {
                grad(L, L.x[L.cur].p, L.grad.p, L.tau.p, kind, false, true);
                restrict_to(L, L.x[L.cur].p, *coarse[i + 1], kind);
            }
        }
        EGGMI_HIP(hipStreamSynchronize(stream));
    } catch (...) {
        drop();
        counting = count0;
        last_launches = launches0;
        throw;
    }
    counting = count0;
    last_launches = launches0;
    release_graphs();
    return ms;
}

// ------------------------------------------------------------------ self-check

double Level::multi_selfcheck(int k, unsigned seed) const
{
    if (ms.k < k || k < 1) throw InvalidArgument("multi_selfcheck: tables of depth >= k not built");
    const Geo geo(lay);
    const long long ns = lay.n_slots();
    const long long total_occ = (long long)st_gc.size();
    const int G = ms.G, R = ms.rounds;
    // a nonlinear update that reads every stencil node in order
    auto mix = [](double acc, int i, double x, double y) { return acc + 0.01 * (i + 1) * std::sin(x + 0.3 * y + i); };
    std::vector<double> X(2 * ns);
    unsigned st = seed * 2654435761u + 12345u;
    for (auto& v : X) {
        st = st * 1664525u + 1013904223u;
        v = (double)(st >> 8) / (double)(1u << 24);
    }
    // replicas consistent with their roots, as after a halo exchange
    for (long long r = 0; r < ns; ++r)
        for (int q = rp_off.empty() ? 0 : rp_off[r]; q < (rp_off.empty() ? 0 : rp_off[r + 1]); ++q)
            for (int c = 0; c < 2; ++c) X[c * ns + rp_dst[q]] = X[c * ns + r];
    auto push = [&](std::vector<double>& Y, long long node) {
        if (cp_off.empty()) return;
        for (int q = cp_off[node]; q < cp_off[node + 1]; ++q)
            for (int c = 0; c < 2; ++c) Y[c * ns + cp_dst[q]] = Y[c * ns + node];
    };
    std::vector<int> st_of(ns, -1);
    for (size_t d = 0; d < st_dof.size(); ++d) st_of[st_dof[d]] = (int)d;

    // ---- k plain synchronous sweeps over the level's own stencils
    std::vector<double> A = X;
    for (int s = 0; s < k; ++s) {
        std::vector<double> B = A;
        for (long long c = 0; c < ns; ++c) {
            if (!mask[c] && st_of[c] < 0) continue;
            double v[2] = {A[c], A[ns + c]};
            int i = 0;
            auto f = [&](long long q) {
                v[0] = mix(v[0], i, A[q], A[ns + q]);
                v[1] = mix(v[1], i, A[ns + q], A[q]);
                ++i;
            };
            if (mask[c]) {
                const long long rs = lay.stride[geo.block(c)][0];
                for (int di = -1; di <= 1; ++di)
                    for (int dj = -1; dj <= 1; ++dj) f(c + di * rs + dj);
            } else {
                const int d = st_of[c];
                for (int p = st_off[d]; p < st_off[d + 1]; ++p) {
                    f(st_gc[p]);
                    for (int a = 0; a < 2; ++a) f(st_gn[(size_t)a * total_occ + p]);
                }
            }
            B[c] = v[0];
            B[ns + c] = v[1];
        }
        for (long long c = 0; c < ns; ++c)
            if (mask[c] || st_of[c] >= 0) push(B, c);
        A.swap(B);
    }

    // ---- the tables' schedule
    std::vector<double> O = X;
    const int S = ms.lds_stride;
    std::vector<double> L[2] = {std::vector<double>(2 * S), std::vector<double>(2 * S)};
    for (const MsTile& T : ms.tiles) {
        for (int ri = 0; ri < T.nreg; ++ri) {
            const MsRegion& Rg = ms.regions[T.reg0 + ri];
            for (int e = 0; e < Rg.h * Rg.w; ++e) {
                const long long s = Rg.base + (long long)(e / Rg.w) * Rg.st0 + e % Rg.w;
                for (int c = 0; c < 2; ++c) L[0][c * S + Rg.lds + e] = L[1][c * S + Rg.lds + e] = X[c * ns + s];
            }
        }
        for (int e = 0; e < T.next; ++e) {
            const long long s = ms.ext[T.ext0 + e];
            const int li = T.n_lds - T.next + e;
            for (int c = 0; c < 2; ++c) L[0][c * S + li] = L[1][c * S + li] = X[c * ns + s];
        }
        for (int s = 0; s < k; ++s) {
            const int r = k - 1 - s;
            const bool fin = r == 0;
            const std::vector<double>& src = L[s & 1];
            std::vector<double>& dst = L[(s + 1) & 1];
            auto emit = [&](int li, long long slot, const double* v) {
                if (fin) {
                    O[slot] = v[0];
                    O[ns + slot] = v[1];
                    push(O, slot);
                } else {
                    dst[li] = v[0];
                    dst[S + li] = v[1];
                }
            };
            for (int it = 0; it < T.syn_cnt[r]; ++it) {
                const int2 en = ms.syn[T.syn0 + it];
                const int li = en.x & 0xFFFF, w = en.x >> 16;
                double v[2] = {src[li], src[S + li]};
                int i = 0;
                for (int di = -1; di <= 1; ++di)
                    for (int dj = -1; dj <= 1; ++dj) {
                        const int q = li + di * w + dj;
                        v[0] = mix(v[0], i, src[q], src[S + q]);
                        v[1] = mix(v[1], i, src[S + q], src[q]);
                        ++i;
                    }
                emit(li, en.y, v);
            }
            for (int it = 0; it < T.st_cnt[r]; ++it) {
                const long long d = T.st0 + it;
                const int li = ms.st_lds[d];
                double v[2] = {src[li], src[S + li]};
                int i = 0;
                for (int o = 0; o < R * G; ++o) {
                    const long long q = (d * R + o / G) * G + o % G;
                    if (ms.meta[q] & kStoredPad) break;
                    const int nodes[3] = {ms.gc[q], ms.gn[q], ms.gn[ms.total + q]};
                    for (int x : nodes) {
                        v[0] = mix(v[0], i, src[x], src[S + x]);
                        v[1] = mix(v[1], i, src[S + x], src[x]);
                        ++i;
                    }
                }
                emit(li, ms.st_slot[d], v);
            }
            if (!fin)
                for (int p = 0; p < T.npush; ++p) {
                    const int ps = ms.push_src[T.push0 + p], pd = ms.push_dst[T.push0 + p];
                    dst[pd] = dst[ps];
                    dst[S + pd] = dst[S + ps];
                }
        }
    }
    // every slot a sweep writes (DOFs and their pushed replicas) must agree
    double err = 0.0;
    for (long long c = 0; c < ns; ++c)
        for (int q = 0; q < 2; ++q) {
            const double x = A[q * ns + c], y = O[q * ns + c];
            if (!(x == y)) err = std::max(err, std::isfinite(x - y) ? std::fabs(x - y) : INFINITY);
        }
    return err;
}

}  // namespace eggmi
