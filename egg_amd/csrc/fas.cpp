// FAS nonlinear multigrid for the structured smoother: the coarse levels
// (reference src/mg_level.hpp:130-455 and src/mg_topology.hpp:121-304 fix
// what a coarse level holds) and the V-cycle driver (src/fas.hpp:251-602),
// over the HIP kernels in mg_kernels.hip / sweep*.hip.
//
// Every coarse level is a Level like the fine one: per-block padded layout,
// its own ghost + share copy table (pushed by the sweep kernel), and DOFs
// that are all synthesized (interior nodes and guarded interface face nodes;
// the face nodes read the level's ghost plane). Transfers run per level
// pair in one launch each.
#include "mg.hpp"
#include "mlp_host.hpp"
#include "session.hpp"

#include <algorithm>
#include <cmath>

namespace eggmi
{

namespace
{
constexpr int kMinCoarsenNodes = 5;
// the level autotuner times one-workgroup smoothing on levels of at most
// this many tiles (it walks them in turn)
constexpr size_t kOneWgMaxTiles = 64;  // an axis halves iff its node count is odd and >= 5 (src/mg_level.hpp:46,142)
constexpr int kMaxHierarchyDepth = 32;

using Idx = std::array<int, 3>;

// One coarsening step: the parent level's layout, relaxed-DOF masks and copy
// tables in; the child's layout, masks and copy tables out. Every parent
// lookup is a dense array over the parent's node slots.
//
// What a child holds (the reference's rules):
// * shapes: per block and axis, halve an odd count n >= 5 to (n + 1) / 2,
//   keep it otherwise; no level when nothing halves;
// * interior DOFs: child node I relaxes iff its parent image I * factor is a
//   synthesized interior DOF and every parent node of its full-weighting
//   stencil (offsets -1, 0, 1 along the halved axes) relaxes, so the
//   restriction reads only gradients a sweep wrote;
// * ghost fills: a parent fill of the ghost behind face (axis, side) at
//   tangential position t * factor becomes the child fill at t; its source,
//   one step inward from the contact in the source block, becomes the child
//   node one step inward from the same contact, at the source's tangential
//   position divided by the source block's factors (dropped when it does not
//   divide: a non-conforming remnant). The contact's normal axis in the
//   source block is the one along which the source steps back onto the
//   parent face node's own copy there (found through the parent share table;
//   dropped when no axis or several do);
// * share pairs: an owner -> copy pair survives iff both nodes are parent
//   images of child nodes;
// * interface DOFs: a surviving owner relaxes iff it lies on exactly one face
//   of its block, strictly inside it, and every parent slot its restriction
//   and patch read relaxes: inside the block a relaxed DOF or a copy of one,
//   behind the face a ghost filled from a relaxed node on the parent level
//   whose child image is filled too.
struct Coarsening {
    const Layout& pl;
    const std::vector<uint8_t>& p_free;  // parent slots its sweeps relax
    const std::vector<uint8_t>& p_int;   // parent synthesized interior DOFs
    const HaloHost& ptab;
    const int D;
    std::vector<Idx> factor, cshape;
    Layout cl;
    std::vector<long long> fill_of;  // parent slot -> slot filling it (ghost fills), -1
    std::vector<long long> rep;      // parent slot -> its share owner (itself when none)
    HaloHost ctab;
    std::vector<uint8_t> c_free, c_int;
    std::vector<uint8_t> c_filled;  // child ghost slots with a fill

    Coarsening(const Layout& parent, const std::vector<uint8_t>& free, const std::vector<uint8_t>& interior,
               const HaloHost& tab)
        : pl(parent), p_free(free), p_int(interior), ptab(tab), D(parent.dim)
    {}

    int block_of(const Layout& L, long long slot) const
    {
        return (int)(std::upper_bound(L.base.begin(), L.base.end(), slot) - L.base.begin()) - 1;
    }
    Idx padded_of(const Layout& L, int b, long long slot) const
    {
        Idx p {0, 0, 0};
        long long r = slot - L.base[b];
        for (int k = 0; k < D; ++k) {
            p[k] = (int)(r / L.stride[b][k]);
            r %= L.stride[b][k];
        }
        return p;
    }

    bool shapes()
    {
        const size_t nb = pl.shape.size();
        factor.assign(nb, {1, 1, 1});
        cshape.assign(nb, {1, 1, 1});
        bool any = false;
        for (size_t b = 0; b < nb; ++b)
            for (int k = 0; k < D; ++k) {
                const int n = pl.shape[b][k];
                const bool halve = n % 2 == 1 && n >= kMinCoarsenNodes;
                factor[b][k] = halve ? 2 : 1;
                cshape[b][k] = halve ? (n + 1) / 2 : n;
                any = any || halve;
            }
        if (!any) return false;
        cl.build(D, cshape);
        c_free.assign(cl.n_slots(), 0);
        c_int.assign(cl.n_slots(), 0);
        c_filled.assign(cl.n_slots(), 0);
        fill_of.assign(pl.n_slots(), -1);
        for (size_t e = 0; e < ptab.sb.size(); ++e)
            fill_of[pl.padded_slot(ptab.db[e], ptab.dp[e].data())] = pl.padded_slot(ptab.sb[e], ptab.sp[e].data());
        rep.resize(pl.n_slots());
        for (long long s = 0; s < pl.n_slots(); ++s) rep[s] = s;
        for (size_t p = 0; p < ptab.share_src.size(); ++p) rep[ptab.share_dst[p]] = ptab.share_src[p];
        return true;
    }

    // calls f(I) for every child interior index of block b, row-major
    template <class F> void each_node(int b, F&& f) const
    {
        const Idx& c = cshape[b];
        Idx I {0, 0, 0};
        const long long n = (long long)c[0] * (D > 1 ? c[1] : 1) * (D > 2 ? c[2] : 1);
        for (long long i = 0; i < n; ++i) {
            long long r = i;
            for (int k = D - 1; k >= 0; --k) {
                I[k] = (int)(r % c[k]);
                r /= c[k];
            }
            f(I);
        }
    }

    void interior_dofs()
    {
        for (int b = 0; b < (int)cshape.size(); ++b) {
            // slot offsets of the full-weighting stencil: -1, 0, 1 along halved axes
            std::vector<long long> stencil {0};
            for (int k = 0; k < D; ++k) {
                if (factor[b][k] == 1) continue;
                std::vector<long long> next;
                for (long long o : stencil)
                    for (int s = -1; s <= 1; ++s) next.push_back(o + s * pl.stride[b][k]);
                stencil.swap(next);
            }
            each_node(b, [&](const Idx& I) {
                Idx F {0, 0, 0};
                for (int k = 0; k < D; ++k) F[k] = I[k] * factor[b][k];
                const long long img = pl.interior_slot(b, F.data());
                if (!p_int[img]) return;
                for (long long o : stencil)
                    if (!p_free[img + o]) return;
                const long long c = cl.interior_slot(b, I.data());
                c_free[c] = 1;
                c_int[c] = 1;
            });
        }
    }

    // the single face (axis, side) a padded index sits behind, axis -1 otherwise
    std::pair<int, int> face_of(const Layout& L, int b, const Idx& p) const
    {
        int axis = -1, side = 0, hits = 0;
        for (int k = 0; k < D; ++k)
            if (p[k] == 0 || p[k] == L.shape[b][k] + 1) {
                ++hits;
                axis = k;
                side = p[k] == 0 ? 0 : 1;
            }
        return hits == 1 ? std::make_pair(axis, side) : std::make_pair(-1, 0);
    }

    void ghost_tables()
    {
        for (size_t e = 0; e < ptab.sb.size(); ++e) {
            const int db = ptab.db[e], sb = ptab.sb[e];
            const Idx& dp = ptab.dp[e];
            const auto [axis, side] = face_of(pl, db, dp);
            if (axis < 0) continue;
            // the child ghost whose parent image this is
            Idx cg {0, 0, 0};
            bool image = true;
            for (int k = 0; k < D && image; ++k) {
                if (k == axis) {
                    cg[k] = side == 0 ? 0 : cshape[db][k] + 1;
                } else {
                    const int t = dp[k] - 1;
                    image = t % factor[db][k] == 0;
                    cg[k] = t / factor[db][k] + 1;
                }
            }
            if (!image) continue;
            // the contact's normal in the source block: the axis along which the
            // source steps back onto a copy of the parent face node
            Idx face = dp;
            face[axis] = side == 0 ? 1 : pl.shape[db][axis];
            const long long face_rep = rep[pl.padded_slot(db, face.data())];
            Idx q = ptab.sp[e];  // padded
            int normal = -1, nside = 0, found = 0;
            for (int k = 0; k < D; ++k) {
                const int n = pl.shape[sb][k];
                for (int s = 0; s < 2; ++s) {
                    const int lq = q[k] - 1;
                    if (lq != (s == 0 ? 1 : n - 2)) continue;
                    Idx tw = q;
                    tw[k] = s == 0 ? 1 : n;  // the face-side node next to the source (padded)
                    if (rep[pl.padded_slot(sb, tw.data())] != face_rep) continue;
                    ++found;
                    normal = k;
                    nside = s;
                }
            }
            if (found != 1) continue;
            Idx cs {0, 0, 0};
            bool conform = true;
            for (int k = 0; k < D && conform; ++k) {
                if (k == normal) {
                    cs[k] = (nside == 0 ? 1 : cshape[sb][k] - 2) + 1;
                } else {
                    const int lq = q[k] - 1;
                    conform = lq % factor[sb][k] == 0;
                    cs[k] = lq / factor[sb][k] + 1;
                }
            }
            if (!conform) continue;
            const long long dst = cl.padded_slot(db, cg.data());
            if (c_filled[dst]) continue;
            c_filled[dst] = 1;
            ctab.sb.push_back(sb);
            ctab.sp.push_back(cs);
            ctab.db.push_back(db);
            ctab.dp.push_back(cg);
        }
    }

    // child interior slot of a parent node slot, -1 when it is no parent image
    long long child_of(long long slot) const
    {
        const int b = block_of(pl, slot);
        const Idx p = padded_of(pl, b, slot);
        Idx c {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            const int l = p[k] - 1;
            if (l < 0 || l >= pl.shape[b][k] || l % factor[b][k] != 0) return -1;
            c[k] = l / factor[b][k];
        }
        return cl.interior_slot(b, c.data());
    }

    void share_pairs()
    {
        for (size_t p = 0; p < ptab.share_src.size(); ++p) {
            const long long so = child_of(ptab.share_src[p]), co = child_of(ptab.share_dst[p]);
            if (so < 0 || co < 0) continue;
            ctab.share_src.push_back(so);
            ctab.share_dst.push_back(co);
        }
    }

    bool readable(long long slot) const { return p_free[slot] || p_free[rep[slot]]; }

    void interface_dofs()
    {
        std::vector<uint8_t> tried(cl.n_slots(), 0);
        for (long long node : ctab.share_src) {
            if (tried[node]) continue;
            tried[node] = 1;
            const int b = block_of(cl, node);
            const Idx cp = padded_of(cl, b, node);
            Idx I {0, 0, 0};
            int axis = -1, side = 0, faces = 0;
            bool inside = true;
            for (int k = 0; k < D; ++k) {
                I[k] = cp[k] - 1;
                if (I[k] < 0 || I[k] >= cshape[b][k]) inside = false;
                if (I[k] == 0 || I[k] == cshape[b][k] - 1) {
                    ++faces;
                    axis = k;
                    side = I[k] == 0 ? 0 : 1;
                }
            }
            if (!inside || faces != 1) continue;
            bool ok = true;
            for (int k = 0; k < D && ok; ++k)
                if (k != axis && (I[k] < 1 || I[k] > cshape[b][k] - 2)) ok = false;
            // the 3^D parent neighbourhood of the image, and the child's
            int o[3] = {-1, -1, -1};
            const int n_off = D == 2 ? 9 : 27;
            for (int e = 0; e < n_off && ok; ++e) {
                int r = e;
                for (int k = 0; k < D; ++k) {
                    o[k] = r % 3 - 1;
                    r /= 3;
                }
                Idx N {0, 0, 0}, C {0, 0, 0};
                for (int k = 0; k < D; ++k) {
                    N[k] = I[k] * factor[b][k] + o[k];
                    C[k] = I[k] + o[k];
                }
                if (N[axis] >= 0 && N[axis] < pl.shape[b][axis]) {
                    ok = readable(pl.interior_slot(b, N.data()));
                    continue;
                }
                Idx gp {0, 0, 0}, cg {0, 0, 0};
                for (int k = 0; k < D; ++k) {
                    gp[k] = k == axis ? (side == 0 ? 0 : pl.shape[b][k] + 1) : N[k] + 1;
                    cg[k] = k == axis ? (side == 0 ? 0 : cshape[b][k] + 1) : C[k] + 1;
                }
                const long long src = fill_of[pl.padded_slot(b, gp.data())];
                ok = src >= 0 && p_free[src] && c_filled[cl.padded_slot(b, cg.data())];
            }
            if (ok) c_free[node] = 1;  // relaxed, but not a synthesized interior DOF of the restriction
        }
    }
};

}  // namespace

// ------------------------------------------------------------------ hierarchy

void Session::build_hierarchy()
{
    if (built_levels >= 0) return;
    built_levels = 0;
    const Layout* pl = &fine.lay;
    const std::vector<uint8_t>* p_free = &fine_free;
    const std::vector<uint8_t>* p_int = &fine.mask;
    const HaloHost* ptab = &fine_halo;
    std::vector<HaloHost> tabs;  // every level's tables (the next step reads its parent's)
    tabs.reserve(kMaxHierarchyDepth);
    for (int lvl = 1; lvl < kMaxHierarchyDepth; ++lvl) {
        Coarsening c(*pl, *p_free, *p_int, *ptab);
        if (!c.shapes()) break;
        c.interior_dofs();
        c.ghost_tables();
        c.share_pairs();
        if (!c.ctab.share_src.empty()) c.interface_dofs();
        const long long n_free = std::count(c.c_free.begin(), c.c_free.end(), (uint8_t)1);
        if (n_free == 0) break;
        auto L = std::make_unique<Level>();
        L->lay = c.cl;
        L->mask = c.c_free;  // every coarse DOF is synthesized
        L->n_synth = n_free;
        L->st_off.assign(1, 0);
        for (size_t e = 0; e < c.ctab.sb.size(); ++e) {
            L->copy_src.push_back((int)L->lay.padded_slot(c.ctab.sb[e], c.ctab.sp[e].data()));
            L->copy_dst.push_back((int)L->lay.padded_slot(c.ctab.db[e], c.ctab.dp[e].data()));
        }
        L->n_halo = (long long)c.ctab.sb.size();
        for (size_t p = 0; p < c.ctab.share_src.size(); ++p) {
            L->copy_src.push_back((int)c.ctab.share_src[p]);
            L->copy_dst.push_back((int)c.ctab.share_dst[p]);
        }
        L->n_share = (long long)c.ctab.share_src.size();
        L->factor = c.factor;
        L->free_mask = c.c_free;
        L->free_interior = c.c_int;
        L->build_push();
        coarse.push_back(std::move(L));
        tabs.push_back(std::move(c.ctab));
        const Level& C = *coarse.back();
        pl = &C.lay;
        p_free = &C.free_mask;
        p_int = &C.free_interior;
        ptab = &tabs.back();
    }
    built_levels = (int)coarse.size();
}

std::vector<std::vector<std::array<int, 3>>> Session::mg_level_shapes(int)
{
    build_hierarchy();
    std::vector<std::vector<std::array<int, 3>>> out;
    for (auto& L : coarse) out.push_back(L->lay.shape);
    return out;
}

// Device residency of coarse level i (lazily, first run_fas that cycles it).
void Session::upload_level(size_t i)
{
    Level& L = *coarse[i];
    if (L.x[0].n) return;
    L.plane = (L.lay.n_slots() + 63) / 64 * 64;
    L.upload_tables();
    apply_level_variant(i);
    const size_t nd = (size_t)L.plane * dim;
    for (auto* a : {&L.x[0], &L.x[1], &L.rx, &L.tau, &L.grad}) {
        a->alloc(nd);
        EGGMI_HIP(hipMemset(a->p, 0, nd * sizeof(float)));
    }
    L.d_free.upload(L.free_mask);
    // transfer tables against the parent level
    std::vector<int> fac;
    for (auto& f : L.factor) fac.insert(fac.end(), {f[0], f[1], f[2]});
    L.d_factor.upload(fac);
    const size_t nb = L.lay.shape.size();
    std::vector<long long> coff(nb + 1, 0);
    for (size_t b = 0; b < nb; ++b) {
        long long c = 1;
        for (int k = 0; k < dim; ++k) c *= L.lay.shape[b][k];
        coff[b + 1] = coff[b] + c;
    }
    L.d_coarse_off.upload(coff);
    L.n_coarse_nodes = coff.back();
}

TransferArgs Session::transfer(const Level& parent, const Level& child, bool prolong) const
{
    TransferArgs t;
    t.fine = parent.blocks.p;
    t.coarse = child.blocks.p;
    t.factor = child.d_factor.p;
    t.nb = (int)child.lay.shape.size();
    t.fplane = parent.plane;
    t.cplane = child.plane;
    if (prolong) {
        t.off = parent.d_fine_off.p;
        t.n = parent.n_fine_nodes;
    } else {
        t.off = child.d_coarse_off.p;
        t.n = child.n_coarse_nodes;
    }
    return t;
}

// ------------------------------------------------------------------ probe

// The transfer operators between coarse level `level` and its parent
// applied to caller fields ([dim][n_slots] doubles of the level they live
// on), for tests of their identities: 0 injection (parent -> child, every
// interior child node), 1 full weighting of a ghost-inclusive parent field
// (child relaxed DOFs, zero elsewhere), 2 prolongation of a child field
// taken as the coarse change (parent relaxed nodes, zero at the other
// interior nodes), 3 (level 1 only) the tau-shifted coarse gradient at the
// injected fine state next to the full-weighted fine gradient, which equal
// each other by construction of tau (two child fields).
void Session::mg_probe(int level, int op, const double* in, double* out)
{
    need_device();
    if (level < 1) throw InvalidArgument("mg probe: level must be >= 1");
    if (op < 0 || op > 3) throw InvalidArgument("mg probe: op must be 0..3");
    if (op == 3 && level != 1) throw InvalidArgument("mg probe: the tau check runs on level 1");
    const size_t n = prepare_fas(level + 1);
    if ((size_t)level > n)
        throw InvalidArgument("mg probe: the hierarchy has " + std::to_string(coarse.size()) + " coarse levels");
    Level& P = level == 1 ? fine : *coarse[level - 2];
    Level& C = *coarse[level - 1];
    auto to_device = [&](const Level& L, const double* src) {
        std::vector<float> h((size_t)L.plane * dim, 0.0f);
        const long long ns = L.lay.n_slots();
        if (src)
            for (int k = 0; k < dim; ++k)
                for (long long q = 0; q < ns; ++q) h[(size_t)k * L.plane + q] = (float)src[(size_t)k * ns + q];
        DArr<float> d;
        d.upload(h);
        return d;
    };
    auto to_host = [&](const Level& L, const float* dev, double* dst) {
        std::vector<float> h((size_t)L.plane * dim);
        EGGMI_HIP(hipStreamSynchronize(stream));
        EGGMI_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        const long long ns = L.lay.n_slots();
        for (int k = 0; k < dim; ++k)
            for (long long q = 0; q < ns; ++q) dst[(size_t)k * ns + q] = h[(size_t)k * L.plane + q];
    };
    if (op == 0 || op == 1) {
        DArr<float> f = to_device(P, in), c = to_device(C, nullptr);
        const TransferArgs t = transfer(P, C, false);
        if (op == 0)
            launch_restrict(dim, t, f.p, c.p, nullptr, nullptr, C.d_free.p, stream);
        else
            launch_restrict(dim, t, nullptr, nullptr, f.p, c.p, C.d_free.p, stream);
        to_host(C, c.p, out);
    } else if (op == 2) {
        DArr<float> c = to_device(C, in), zero = to_device(C, nullptr), f = to_device(P, nullptr);
        launch_prolong(dim, transfer(P, C, true), c.p, zero.p, f.p, P.d_free.p, false, nullptr, nullptr, stream);
        to_host(P, f.p, out);
    } else {
        float* x = P.x[P.cur].p;
        grad(P, x, P.grad.p, nullptr, 0, false, true);
        restrict_to(P, x, C, 0);
        DArr<float> shifted = to_device(C, nullptr), rg = to_device(C, nullptr);
        grad(C, C.rx.p, shifted.p, C.tau.p, 0);
        launch_restrict(dim, transfer(P, C, false), nullptr, nullptr, P.grad.p, rg.p, C.d_free.p, stream);
        to_host(C, shifted.p, out);
        to_host(C, rg.p, out + (size_t)dim * C.lay.n_slots());
    }
    EGGMI_HIP(hipStreamSynchronize(stream));
}

// ------------------------------------------------------------------ driver

void Session::smooth(Level& L, int sweeps, int kind, float omega, const float* tau)
{
    // the learned proposal, when set, runs in the fine level's (unshifted)
    // smoothing sweeps as in plain runs; the other Jacobi buffer holds each
    // DOF's state before the last sweep (or correction), its displacement input
    const bool propose = proposal != nullptr && &L == &fine && tau == nullptr && kind != 1;
    if (L.one_wg && tau && !propose && !multi_k(L, propose)) {
        if (sweeps <= 0) return;
        float* x0 = L.x[L.cur].p;
        float* x1 = L.x[1 - L.cur].p;
        SweepArgs a = L.sweep_args(x0, x1, omega, 0.0f, tau);
        a.counters = counters_for(L);
        launch_level_smooth(dim, kind, L.variant, a, x0, x1, sweeps, stream);
        if (sweeps % 2) L.cur = 1 - L.cur;
        ++last_launches;
        return;
    }
    if (const int mk = multi_k(L, propose)) {  // k sweeps per launch
        for (int s = 0; s < sweeps;) {
            const int kk = std::min(mk, sweeps - s);
            launch_multi(L, L.x[L.cur].p, L.x[1 - L.cur].p, kk, kind, omega, 0.0f, tau);
            L.cur = 1 - L.cur;
            ++last_launches;
            s += kk;
        }
        return;
    }
    for (int s = 0; s < sweeps; ++s) {
        float* xin = L.x[L.cur].p;
        float* xout = L.x[1 - L.cur].p;
        SweepArgs a = L.sweep_args(xin, xout, omega, 0.0f, tau);
        a.counters = counters_for(L);
        const bool here = propose && (proposal_every == 1 || s == 0);
        if (here) {
            a.mlp = proposal->view();
            a.prop_mode = proposal_mode;
            a.prop_scale = proposal_scale;
        }
        launch_sweep(dim, kind, here ? proposal_variant() : L.variant, a, stream);
        L.cur = 1 - L.cur;
        ++last_launches;
    }
}

void Session::grad(Level& L, const float* x, float* out, const float* tau, int kind, bool sub, bool push)
{
    SweepArgs a = L.sweep_args(x, nullptr, 1.0f, 0.0f, tau);
    a.gout = out;
    a.gsub = sub;
    a.gpush = push;
    launch_grad(dim, kind, L.variant, a, stream);
}

// Restrict the parent state x_parent into child (x, rx and the second
// Jacobi buffer) and the parent's (shifted) gradient, which parent.grad holds
// with its replicas (a pushed gradient launch), into tau = grad_c(Rx) - R grad
// (at the child's DOF slots, the only ones its sweeps read).
void Session::restrict_to(Level& parent, const float* x_parent, Level& child, int kind)
{
    const TransferArgs t = transfer(parent, child, false);
    // one launch writes both Jacobi buffers and rx, replica slots included
    RestrictPush push;
    push.x1 = child.x[1].p;
    push.rx = child.rx.p;
    push.off = child.d_rp_off.p;
    push.dst = child.d_rp_dst.p;
    push.is_dst = child.d_is_dst.p;
    launch_restrict(dim, t, x_parent, child.x[0].p, parent.grad.p, child.tau.p, child.d_free.p, stream, push);
    child.cur = 0;
    grad(child, child.rx.p, child.tau.p, nullptr, kind, true);  // tau = grad_c(Rx) - R grad at the DOFs
}

void Session::coarse_vcycle(size_t li, size_t n_levels, int kind, const FasParams& o)
{
    Level& L = *coarse[li];
    if (li + 1 == n_levels) {
        smooth(L, o.nu_coarse, kind, o.omega, L.tau.p);
        return;
    }
    Level& N = *coarse[li + 1];
    smooth(L, o.nu_pre, kind, o.omega, L.tau.p);
    grad(L, L.x[L.cur].p, L.grad.p, L.tau.p, kind, false, true);
    restrict_to(L, L.x[L.cur].p, N, kind);
    coarse_vcycle(li + 1, n_levels, kind, o);
    // x += P(x_c - rx_c) at free slots and their replicas, one launch (the
    // replicas of fixed nodes keep the restriction's values)
    launch_prolong(dim, transfer(L, N, true), N.x[N.cur].p, N.rx.p, L.x[L.cur].p, L.d_free.p, true,
                   L.n_push ? L.d_cp_off.p : nullptr, L.n_push ? L.d_cp_dst.p : nullptr, stream);
    smooth(L, o.nu_post, kind, o.omega, L.tau.p);
}

std::pair<std::vector<double>, std::vector<double>> Session::run_fas(int n_cycles, int nu_pre, int nu_post,
                                                                     int nu_coarse, double omega, int max_levels,
                                                                     int kind)
{
    need_device();
    if (n_cycles <= 0) return {{}, {}};
    const int per_cycle = nu_pre + nu_post;
    const size_t n_coarse = prepare_fas(max_levels);
    if (n_coarse == 0) return run_jacobi(n_cycles * per_cycle, kind, 0.0, omega, per_cycle);
    if (weighted)
        throw InvalidArgument("run_fas: per-sample energy weights are not supported by the multigrid path "
                              "(its coarse levels and line search use the unweighted cell energy)");
    if (!fine.structured_energy)
        throw InvalidArgument("run_fas: the line search needs the structured cell energy stencil (identity "
                              "W_inv on the energy stencil)");

    Level& F = fine;
    FasParams o;
    o.nu_pre = nu_pre;
    o.nu_post = nu_post;
    o.omega = (float)omega;
    {
        int ext = 0;
        for (auto& sh : coarse[n_coarse - 1]->lay.shape)
            for (int k = 0; k < dim; ++k) ext = std::max(ext, sh[k]);
        o.nu_coarse = std::min(nu_coarse, 2 * ext);
    }
    // reports (a buffer of FAS's own, which captured cycles append to):
    // energies at [0, cap), min dets at [cap, 2 cap)
    if ((long long)d_fas_reports.n < 2LL * n_cycles) {
        release_graphs();
        d_fas_reports.alloc(2 * std::max(n_cycles, 64));
    }
    const int cap = (int)(d_fas_reports.n / 2);
    if (!d_ls.n) d_ls.alloc(4);
    last_launches = 0;
    last_fas_graph_kernels = 0;
    counters_begin();
    const bool graphs_on = graph_chunk > 0 && !counting;
    if (graphs_on) {
        if (!d_ladder_parts.n) {
            d_ladder_parts.alloc((size_t)2 * (kLadder + 1) * kReduceBlocks);  // cells + directional
            d_ladder_out.alloc(2 * (kLadder + 1));
            d_cyc_rep.alloc(2);
            d_alpha.alloc(1);
            d_rep_count.alloc(1);
        }
        const int zero = 0;
        EGGMI_HIP(hipMemcpyAsync(d_rep_count.p, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
        EGGMI_HIP(hipStreamSynchronize(stream));  // the source is a host stack value
    }
    EGGMI_HIP(hipEventRecord(ev0, stream));
    // consistent halo on the resident state, both fine Jacobi buffers equal
    halo(F, F.x[F.cur].p);
    F.next_run(stream);
    EGGMI_HIP(hipMemcpyAsync(F.x[1 - F.cur].p, F.x[F.cur].p, (size_t)F.plane * dim * sizeof(float),
                             hipMemcpyDeviceToDevice, stream));
    pair_valid = true;
    const float tol = 1e-6f;  // FP32 accept slack (reference src/core.hpp:40)
    for (int cyc = 0; cyc < n_cycles; ++cyc) {
        if (graphs_on) {
            const int parity = F.x[F.cur].p == F.x[0].p ? 0 : 1;
            EGGMI_HIP(hipGraphLaunch(fas_graph(parity, kind, o, n_coarse), stream));
            const FasGraph& g = fas_graphs[parity];
            last_launches += g.launches;
            last_fas_graph_kernels = g.kernels;
            if ((o.nu_pre + o.nu_post) % 2) F.cur = 1 - F.cur;
            continue;
        }
        fas_head(kind, o, n_coarse);
        float* x = F.x[F.cur].p;
        // safeguarded application (reference src/fas.hpp:549-581), decided
        // on the host one trial pair at a time
        const CellReduceArgs ra = fine_cells(x);
        double ls[4];
        auto ladder = [&](float a0, float a1) {
            if (F.structured_energy) {
                launch_linesearch_pair(dim, kind, ra, F.corr.p, a0, a1, partials.p, d_ls.p, stream,
                                       F.es_dir.n ? &F.es_dir : nullptr);
            } else {
                throw Internal("run_fas: the line search needs the structured cell energy stencil");
            }
            EGGMI_HIP(hipMemcpyAsync(ls, d_ls.p, sizeof(ls), hipMemcpyDeviceToHost, stream));
            EGGMI_HIP(hipStreamSynchronize(stream));
        };
        auto ok = [&](double e_pre, double e_t, double m_t) {
            return std::isfinite(e_t) && (e_t <= e_pre + tol) && m_t > 0.0;
        };
        float alpha = 0.0f;
        ladder(0.0f, 1.0f);
        const double e_pre = ls[0];
        if (ok(e_pre, ls[2], ls[3])) alpha = 1.0f;
        const int max_bt = kLadder;
        float a = 0.5f;
        for (int it = 1; alpha == 0.0f && it < max_bt; it += 2, a *= 0.25f) {
            ladder(a, a * 0.5f);
            if (ok(e_pre, ls[0], ls[1]))
                alpha = a;
            else if (it + 1 < max_bt && ok(e_pre, ls[2], ls[3]))
                alpha = a * 0.5f;
        }
        if (alpha != 0.0f) {
            // in place at free slots; the copy refreshes their replicas (the
            // other Jacobi buffer's DOF slots are rewritten by the next sweep)
            launch_masked_axpy(x, F.corr.p, alpha, F.d_free.p, F.lay.n_slots(), F.plane, dim, stream);
            halo(F, x);
        }
        fas_tail(kind, o, d_fas_reports.p + cyc, d_fas_reports.p + cap + cyc);
    }
    EGGMI_HIP(hipEventRecord(ev1, stream));
    std::vector<double> e(n_cycles), m(n_cycles);
    EGGMI_HIP(hipMemcpyAsync(e.data(), d_fas_reports.p, sizeof(double) * n_cycles, hipMemcpyDeviceToHost, stream));
    EGGMI_HIP(hipMemcpyAsync(m.data(), d_fas_reports.p + cap, sizeof(double) * n_cycles, hipMemcpyDeviceToHost,
                             stream));
    EGGMI_HIP(hipStreamSynchronize(stream));
    counters_end();
    float ms = 0;
    EGGMI_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    last_ms = ms;
    return {e, m};
}

size_t Session::prepare_fas(int max_levels)
{
    build_hierarchy();
    const size_t n_coarse = max_levels > 1 ? std::min(coarse.size(), (size_t)(max_levels - 1)) : 0;
    if (n_coarse == 0) return 0;
    Level& F = fine;
    if (!F.grad.n) {
        const size_t nd = (size_t)F.plane * dim;
        for (auto* a : {&F.grad, &F.corr}) {
            a->alloc(nd);
            EGGMI_HIP(hipMemset(a->p, 0, nd * sizeof(float)));
        }
        F.d_free.upload(fine_free);
    }
    for (size_t i = 0; i < n_coarse; ++i) upload_level(i);
    // prolongation enumerates the interior nodes of the parent of each level
    for (size_t i = 0; i < n_coarse; ++i) {
        Level& P = i == 0 ? F : *coarse[i - 1];
        if (P.d_fine_off.n) continue;
        const size_t nb = P.lay.shape.size();
        std::vector<long long> off(nb + 1, 0);
        for (size_t b = 0; b < nb; ++b) {
            long long c = 1;
            for (int k = 0; k < dim; ++k) c *= P.lay.shape[b][k];
            off[b + 1] = off[b] + c;
        }
        P.d_fine_off.upload(off);
        P.n_fine_nodes = off.back();
    }
    return n_coarse;
}

// Coarse levels run below one wave per SIMD, so a sweep there lasts about
// one wave's lifetime and the tile shape that spreads a DOF over the most
// lanes can win, while the fine level is throughput-bound. Each level is
// timed from the state a cycle's first restriction gives it (the resident
// fine state restricted down the chain, no smoothing); coarse-level buffers
// are rewritten by every cycle's restriction, and the fine state is only
// read (its replicas refreshed, as run_fas does first).
std::vector<double> Session::autotune_levels(int n_sweeps, double omega, int max_levels, int kind)
{
    need_device();
    if (n_sweeps < 1) throw InvalidArgument("autotune_levels: n_sweeps must be >= 1");
    if (kind == 1) throw InvalidArgument("autotune_levels: FAS smooths the barrier phases only");
    const size_t n_coarse = prepare_fas(max_levels);
    const int nt = num_tile_variants(dim), nv = 2 * nt;  // per-sweep launches, then one workgroup
    std::vector<double> ms(n_coarse * nv, -1.0);
    if (n_coarse == 0) return ms;
    const bool count0 = counting;
    const long long launches0 = last_launches;
    counting = false;
    std::vector<int> picks(n_coarse);
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
            for (int v = 0; v < nv; ++v) {
                // timed as the captured graph a cycle replays them in: the
                // launch gaps are the graph's, not the host's launch rate
                L.make_tiles(v % nt);
                L.one_wg = v >= nt;
                // one workgroup walks every tile in turn: only small levels
                // of synthesized DOFs can gain from it (-1: not timed)
                if (L.one_wg && (!L.st_dof.empty() || !L.con_tag.empty() || L.tiles.size() > kOneWgMaxTiles))
                    continue;
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
                EGGMI_HIP(hipGraphLaunch(exec, stream));  // warm the variant's code
                reset();
                EGGMI_HIP(hipEventRecord(ev0, stream));
                EGGMI_HIP(hipGraphLaunch(exec, stream));
                EGGMI_HIP(hipEventRecord(ev1, stream));
                EGGMI_HIP(hipEventSynchronize(ev1));
                float t = 0;
                EGGMI_HIP(hipEventElapsedTime(&t, ev0, ev1));
                ms[i * nv + v] = t;
                drop();
                reset();
            }
            int best = 0;
            for (int v = 1; v < nv; ++v)
                if (ms[i * nv + v] >= 0.0 && ms[i * nv + v] < ms[i * nv + best]) best = v;
            picks[i] = best;
            L.make_tiles(best % nt);
            L.one_wg = best >= nt;
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
        for (size_t i = 0; i < n_coarse; ++i) apply_level_variant(i);
        throw;
    }
    counting = count0;
    last_launches = launches0;
    std::vector<int> lv = level_variant;
    if (lv.size() < n_coarse) lv.resize(n_coarse, -1);
    for (size_t i = 0; i < n_coarse; ++i) lv[i] = picks[i];
    set_level_variants(lv);
    return ms;
}

CellReduceArgs Session::fine_cells(const float* x) const
{
    const Level& F = fine;
    CellReduceArgs ra;
    ra.x = x;
    ra.plane = F.plane;
    ra.blocks = F.blocks.p;
    ra.cell_off = F.d_cell_off.p;
    ra.nblocks = (int)F.lay.shape.size();
    ra.ncells = F.ncells;
    return ra;
}

void Session::fas_head(int kind, const FasParams& o, size_t n_coarse)
{
    Level& F = fine;
    Level& C0 = *coarse[0];
    smooth(F, o.nu_pre, kind, o.omega, nullptr);
    float* x = F.x[F.cur].p;
    grad(F, x, F.grad.p, nullptr, kind, false, true);
    restrict_to(F, x, C0, kind);
    coarse_vcycle(0, n_coarse, kind, o);
    // fine correction P(x_c - rx_c), zero where not free, pushed into the
    // replicas of the free slots (the others' replicas hold zero throughout)
    launch_prolong(dim, transfer(F, C0, true), C0.x[C0.cur].p, C0.rx.p, F.corr.p, F.d_free.p, false,
                   F.n_push ? F.d_cp_off.p : nullptr, F.n_push ? F.d_cp_dst.p : nullptr, stream, F.d_is_dst.p);
}

void Session::fas_device_ladder(int kind)
{
    Level& F = fine;
    if (!F.structured_energy) throw Internal("run_fas: the line search needs the structured cell energy stencil");
    float* x = F.x[F.cur].p;
    launch_linesearch_ladder(dim, kind, fine_cells(x), F.corr.p, d_ladder_parts.p, (double)1e-6f, d_ladder_out.p,
                             d_alpha.p, stream, F.es_dir.n ? &F.es_dir : nullptr);
    // the step and its replicas (the copies of fixed roots never change)
    launch_masked_axpy_dev(x, F.corr.p, d_alpha.p, F.d_free.p, F.lay.n_slots(), F.plane, dim, stream,
                           F.n_push ? F.d_cp_off.p : nullptr, F.n_push ? F.d_cp_dst.p : nullptr);
}

void Session::fas_tail(int kind, const FasParams& o, double* out_e, double* out_m, double* rep, int cap, int* count)
{
    Level& F = fine;
    smooth(F, o.nu_post, kind, o.omega, nullptr);
    reduce(F, F.x[F.cur].p, kind, 0.0, out_e, out_m, rep, cap, count);
}

hipGraphExec_t Session::fas_graph(int parity, int kind, const FasParams& o, size_t n_coarse)
{
    FasGraph& g = fas_graphs[parity];
    const bool propose = proposal != nullptr && kind != 1;
    const void* mlp = propose ? (const void*)proposal->view().fwd : nullptr;
    if (g.exec && g.kind == kind && g.variant == tile_variant && g.nu_pre == o.nu_pre && g.nu_post == o.nu_post &&
        g.nu_coarse == o.nu_coarse && g.n_coarse == n_coarse && g.omega == o.omega && g.rep == d_fas_reports.p &&
        g.mlp == mlp && (!propose || (g.mode == proposal_mode && g.scale == proposal_scale && g.every == proposal_every)))
        return g.exec;
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = FasGraph {};
    Level& F = fine;
    // the capture enqueues (does not run) one cycle from this parity; the
    // host-side buffer flips it makes are undone afterwards
    const int cur0 = F.cur;
    F.cur = parity;
    std::vector<int> ccur;
    for (auto& c : coarse) ccur.push_back(c->cur);
    const long long launches0 = last_launches;
    hipGraph_t graph = nullptr;
    EGGMI_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
        fas_head(kind, o, n_coarse);
        fas_device_ladder(kind);
        // the cycle's report is appended by the reduction's last launch
        fas_tail(kind, o, d_cyc_rep.p, d_cyc_rep.p + 1, d_fas_reports.p, (int)(d_fas_reports.n / 2), d_rep_count.p);
    } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        F.cur = cur0;
        for (size_t i = 0; i < coarse.size(); ++i) coarse[i]->cur = ccur[i];
        last_launches = launches0;
        throw;
    }
    EGGMI_HIP(hipStreamEndCapture(stream, &graph));
    size_t n_nodes = 0;
    EGGMI_HIP(hipGraphGetNodes(graph, nullptr, &n_nodes));
    std::vector<hipGraphNode_t> nodes(n_nodes);
    if (n_nodes) EGGMI_HIP(hipGraphGetNodes(graph, nodes.data(), &n_nodes));
    long long n_kernels = 0;
    for (auto nd : nodes) {
        hipGraphNodeType t {};
        if (hipGraphNodeGetType(nd, &t) == hipSuccess && t == hipGraphNodeTypeKernel) ++n_kernels;
    }
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    g.launches = last_launches - launches0;
    g.kernels = n_kernels;
    F.cur = cur0;
    for (size_t i = 0; i < coarse.size(); ++i) coarse[i]->cur = ccur[i];
    last_launches = launches0;
    EGGMI_HIP(ie);
    g.kind = kind;
    g.variant = tile_variant;
    g.nu_pre = o.nu_pre;
    g.nu_post = o.nu_post;
    g.nu_coarse = o.nu_coarse;
    g.n_coarse = n_coarse;
    g.omega = o.omega;
    g.rep = d_fas_reports.p;
    g.mlp = mlp;
    g.mode = proposal_mode;
    g.scale = proposal_scale;
    g.every = proposal_every;
    return g.exec;
}

}  // namespace eggmi
