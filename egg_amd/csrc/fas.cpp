// FAS nonlinear multigrid for the structured smoother: hierarchy build
// (reference src/mg_level.hpp:130-455), coarse interface tables
// (src/mg_topology.hpp:121-304) and the V-cycle driver (src/fas.hpp:251-602),
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
#include <map>
#include <unordered_map>
#include <cmath>
#include <unordered_set>

namespace eggmi
{

namespace
{
constexpr int kMinCoarsenNodes = 5;  // reference src/mg_level.hpp:46
constexpr int kMaxHierarchyDepth = 32;

struct Loc {
    int b;
    int p[3];
};

Loc locate(const Layout& L, long long slot)
{
    int b = 0;
    while (b + 1 < (int)L.base.size() && L.base[b + 1] <= slot) ++b;
    long long rem = slot - L.base[b];
    Loc r {b, {0, 0, 0}};
    for (int k = 0; k < L.dim; ++k) {
        r.p[k] = (int)(rem / L.stride[b][k]);
        rem %= L.stride[b][k];
    }
    return r;
}

// (axis, side) of a face ghost padded index, axis -1 when not a face ghost
std::pair<int, int> ghost_face(const Layout& L, int b, const int* p)
{
    int axis = -1, side = 0;
    for (int k = 0; k < L.dim; ++k) {
        const bool lo = p[k] == 0, hi = p[k] == L.shape[b][k] + 1;
        if (lo || hi) {
            if (axis >= 0) return {-1, 0};
            axis = k;
            side = hi ? 1 : 0;
        }
    }
    return {axis, side};
}

// Child ghost + share tables from the parent's (reference src/mg_topology.hpp:121-304).
HaloHost coarsen_halo(const HaloHost& parent, const Layout& pl, const Layout& cl,
                      const std::vector<std::array<int, 3>>& factor)
{
    HaloHost child;
    const int D = pl.dim;
    if (parent.sb.empty() && parent.share_src.empty()) return child;
    std::unordered_map<long long, size_t> by_dst;
    for (size_t e = 0; e < parent.sb.size(); ++e) by_dst.emplace(pl.padded_slot(parent.db[e], parent.dp[e].data()), e);

    struct FaceInfo {
        std::array<int, 3> first {};
        std::array<bool, 3> same {};
        bool seen = false, valid = true;
        int normal = -1, side = 0;
    };
    std::map<std::array<int, 4>, FaceInfo> faces;  // (dst blk, axis, side, src blk)
    for (size_t e = 0; e < parent.sb.size(); ++e) {
        const int db = parent.db[e];
        const auto [axis, side] = ghost_face(pl, db, parent.dp[e].data());
        if (axis < 0) continue;
        FaceInfo& fi = faces[{db, axis, side, parent.sb[e]}];
        std::array<int, 3> logical = parent.sp[e];
        for (int k = 0; k < D; ++k) logical[k] -= 1;
        if (!fi.seen) {
            fi.seen = true;
            fi.first = logical;
            fi.same = {true, true, true};
        } else {
            for (int k = 0; k < D; ++k)
                if (logical[k] != fi.first[k]) fi.same[k] = false;
        }
    }
    for (auto& [key, fi] : faces) {
        const auto& sn = pl.shape[key[3]];
        for (int k = 0; k < D; ++k) {
            if (!fi.same[k]) continue;
            const bool lo = fi.first[k] == 1, hi = fi.first[k] == sn[k] - 2;
            if (!lo && !hi) continue;
            if (fi.normal >= 0) {
                fi.valid = false;
                break;
            }
            fi.normal = k;
            fi.side = hi ? 1 : 0;
        }
        if (fi.normal < 0) fi.valid = false;
    }
    std::map<std::array<int, 3>, bool> face_set;
    for (const auto& [key, fi] : faces) face_set[{key[0], key[1], key[2]}] = true;

    for (const auto& fentry : face_set) {
        const int db = fentry.first[0], axis = fentry.first[1], side = fentry.first[2];
        const auto& cn = cl.shape[db];
        const auto& fn = pl.shape[db];
        long long count = 1;
        int tshape[3] = {1, 1, 1};
        for (int k = 0; k < D; ++k) {
            tshape[k] = (k == axis) ? 1 : cn[k];
            count *= tshape[k];
        }
        for (long long i = 0; i < count; ++i) {
            int T[3] = {0, 0, 0};
            long long rem = i;
            for (int k = D - 1; k >= 0; --k) {
                T[k] = (int)(rem % tshape[k]);
                rem /= tshape[k];
            }
            int fine_ghost[3] = {0, 0, 0}, child_ghost[3] = {0, 0, 0};
            for (int k = 0; k < D; ++k) {
                if (k == axis) {
                    fine_ghost[k] = side == 0 ? 0 : fn[k] + 1;
                    child_ghost[k] = side == 0 ? 0 : cn[k] + 1;
                } else {
                    fine_ghost[k] = T[k] * factor[db][k] + 1;
                    child_ghost[k] = T[k] + 1;
                }
            }
            const auto it = by_dst.find(pl.padded_slot(db, fine_ghost));
            if (it == by_dst.end()) continue;
            const size_t e = it->second;
            const int sb = parent.sb[e];
            const auto fit = faces.find({db, axis, side, sb});
            if (fit == faces.end() || !fit->second.valid) continue;
            const FaceInfo& fi = fit->second;
            const auto& snc = cl.shape[sb];
            const auto& snf = pl.shape[sb];
            std::array<int, 3> q1 = parent.sp[e];
            for (int k = 0; k < D; ++k) q1[k] -= 1;
            if (q1[fi.normal] != (fi.side == 0 ? 1 : snf[fi.normal] - 2)) continue;
            std::array<int, 3> c {0, 0, 0};
            bool ok = true;
            for (int k = 0; k < D && ok; ++k) {
                if (k == fi.normal) {
                    c[k] = fi.side == 0 ? 1 : snc[k] - 2;
                    continue;
                }
                const int f = factor[sb][k];
                if (q1[k] % f != 0) {
                    ok = false;
                    break;
                }
                c[k] = q1[k] / f;
            }
            if (!ok) continue;
            for (int k = 0; k < D; ++k) c[k] += 1;
            child.sb.push_back(sb);
            child.sp.push_back(c);
            child.db.push_back(db);
            child.dp.push_back({child_ghost[0], child_ghost[1], child_ghost[2]});
        }
    }
    auto to_child = [&](long long slot, long long& out) {
        const Loc l = locate(pl, slot);
        int logical[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            if (l.p[k] == 0) return false;
            logical[k] = l.p[k] - 1;
        }
        int c[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            const int f = factor[l.b][k];
            if (logical[k] % f != 0) return false;
            c[k] = logical[k] / f;
            if (c[k] >= cl.shape[l.b][k]) return false;
        }
        out = cl.interior_slot(l.b, c);
        return true;
    };
    for (size_t p = 0; p < parent.share_src.size(); ++p) {
        long long so = 0, co = 0;
        if (to_child(parent.share_src[p], so) && to_child(parent.share_dst[p], co)) {
            child.share_src.push_back(so);
            child.share_dst.push_back(co);
        }
    }
    return child;
}

}  // namespace

// ------------------------------------------------------------------ hierarchy

void Session::build_hierarchy()
{
    if (built_levels >= 0) return;
    built_levels = 0;
    const int D = dim;
    const Layout* pl = &fine.lay;
    std::vector<uint8_t> p_free = fine_free, p_int = fine.mask;
    HaloHost p_tab = fine_halo;
    for (int lvl = 1; lvl < kMaxHierarchyDepth; ++lvl) {
        const size_t nb = pl->shape.size();
        std::vector<std::array<int, 3>> factor(nb, {1, 1, 1}), cshape(nb, {1, 1, 1});
        bool any = false;
        for (size_t b = 0; b < nb; ++b)
            for (int k = 0; k < D; ++k) {
                const int n = pl->shape[b][k];
                const bool halve = (n % 2 == 1) && n >= kMinCoarsenNodes;
                factor[b][k] = halve ? 2 : 1;
                cshape[b][k] = halve ? (n + 1) / 2 : n;
                any = any || halve;
            }
        if (!any) break;
        auto L = std::make_unique<Level>();
        L->lay.build(D, cshape);
        const long long ns = L->lay.n_slots();
        std::vector<uint8_t> c_free(ns, 0), c_int(ns, 0);
        L->mask.assign(ns, 0);
        const int NO = D == 2 ? 9 : 27;
        for (size_t b = 0; b < nb; ++b) {
            const auto& cs = cshape[b];
            const long long cnt = (long long)cs[0] * cs[1] * (D == 3 ? cs[2] : 1);
            for (long long i = 0; i < cnt; ++i) {
                int I[3] = {0, 0, 0};
                long long rem = i;
                for (int k = D - 1; k >= 0; --k) {
                    I[k] = (int)(rem % cs[k]);
                    rem /= cs[k];
                }
                int F[3] = {0, 0, 0};
                for (int k = 0; k < D; ++k) F[k] = I[k] * factor[b][k];
                if (!p_int[pl->interior_slot((int)b, F)]) continue;
                bool ok = true;
                for (int e = 0; e < NO && ok; ++e) {
                    int r = e, P[3] = {F[0], F[1], F[2]};
                    bool valid = true;
                    for (int k = 0; k < D; ++k) {
                        const int o = r % 3 - 1;
                        r /= 3;
                        if (factor[b][k] == 1) {
                            if (o != 0) {
                                valid = false;
                                break;
                            }
                        } else {
                            P[k] = F[k] + o;
                        }
                    }
                    if (!valid) continue;
                    ok = p_free[pl->interior_slot((int)b, P)] != 0;
                }
                if (!ok) continue;
                const long long cid = L->lay.interior_slot((int)b, I);
                c_free[cid] = 1;
                c_int[cid] = 1;
            }
        }
        HaloHost c_tab = coarsen_halo(p_tab, *pl, L->lay, factor);
        if (!c_tab.share_src.empty()) {
            // guarded interface DOFs (reference src/mg_level.hpp:258-392)
            std::unordered_map<long long, size_t> parent_ghost;
            for (size_t e = 0; e < p_tab.sb.size(); ++e) parent_ghost.emplace(pl->padded_slot(p_tab.db[e], p_tab.dp[e].data()), e);
            std::unordered_map<long long, long long> copy_owner;
            for (size_t p = 0; p < p_tab.share_src.size(); ++p) copy_owner.emplace(p_tab.share_dst[p], p_tab.share_src[p]);
            auto read_ok = [&](long long nid) {
                if (p_free[nid]) return true;
                const auto it = copy_owner.find(nid);
                return it != copy_owner.end() && p_free[it->second] != 0;
            };
            std::unordered_set<long long> child_ghost;
            for (size_t e = 0; e < c_tab.sb.size(); ++e) child_ghost.insert(L->lay.padded_slot(c_tab.db[e], c_tab.dp[e].data()));
            std::unordered_set<long long> seen;
            for (long long node : c_tab.share_src) {
                if (!seen.insert(node).second) continue;
                const Loc l = locate(L->lay, node);
                const int b = l.b;
                const auto& cn = L->lay.shape[b];
                const auto& fn = pl->shape[b];
                int I[3] = {0, 0, 0};
                bool interior = true;
                for (int k = 0; k < D; ++k) {
                    if (l.p[k] == 0) interior = false;
                    I[k] = l.p[k] - 1;
                }
                if (!interior) continue;
                int axis = -1, side = 0;
                bool ok = true;
                for (int k = 0; k < D && ok; ++k) {
                    const bool lo = I[k] == 0, hi = I[k] == cn[k] - 1;
                    if (lo || hi) {
                        if (axis >= 0) ok = false;
                        axis = k;
                        side = hi ? 1 : 0;
                    }
                }
                if (!ok || axis < 0) continue;
                for (int k = 0; k < D && ok; ++k)
                    if (k != axis && (I[k] < 1 || I[k] > cn[k] - 2)) ok = false;
                if (!ok) continue;
                for (int e = 0; e < NO && ok; ++e) {
                    int r = e;
                    long long N[3] = {0, 0, 0}, C[3] = {0, 0, 0};
                    for (int k = 0; k < D; ++k) {
                        const int o = r % 3 - 1;
                        r /= 3;
                        N[k] = (long long)I[k] * factor[b][k] + o;
                        C[k] = I[k] + o;
                    }
                    if (N[axis] < 0 || N[axis] >= fn[axis]) {
                        int gp[3] = {0, 0, 0}, cgp[3] = {0, 0, 0};
                        for (int k = 0; k < D; ++k) {
                            if (k == axis) {
                                gp[k] = side == 0 ? 0 : fn[k] + 1;
                                cgp[k] = side == 0 ? 0 : cn[k] + 1;
                            } else {
                                gp[k] = (int)N[k] + 1;
                                cgp[k] = (int)C[k] + 1;
                            }
                        }
                        const auto it = parent_ghost.find(pl->padded_slot(b, gp));
                        ok = it != parent_ghost.end() &&
                             p_free[pl->padded_slot(p_tab.sb[it->second], p_tab.sp[it->second].data())] != 0 &&
                             child_ghost.count(L->lay.padded_slot(b, cgp)) != 0;
                    } else {
                        int P[3] = {(int)N[0], (int)N[1], (int)N[2]};
                        ok = read_ok(pl->interior_slot(b, P));
                    }
                }
                if (!ok) continue;
                c_free[node] = 1;  // free, not free_interior
            }
        }
        long long n_free = 0;
        for (long long s = 0; s < ns; ++s) n_free += c_free[s];
        if (n_free == 0) break;
        L->mask = c_free;  // every coarse DOF is synthesized
        L->n_synth = n_free;
        L->st_off.assign(1, 0);
        for (size_t e = 0; e < c_tab.sb.size(); ++e) {
            L->copy_src.push_back((int)L->lay.padded_slot(c_tab.sb[e], c_tab.sp[e].data()));
            L->copy_dst.push_back((int)L->lay.padded_slot(c_tab.db[e], c_tab.dp[e].data()));
        }
        L->n_halo = (long long)c_tab.sb.size();
        for (size_t p = 0; p < c_tab.share_src.size(); ++p) {
            L->copy_src.push_back((int)c_tab.share_src[p]);
            L->copy_dst.push_back((int)c_tab.share_dst[p]);
        }
        L->n_share = (long long)c_tab.share_src.size();
        L->factor = factor;
        L->free_mask = c_free;
        L->free_interior = c_int;
        L->build_push();
        coarse.push_back(std::move(L));
        Level& C = *coarse.back();
        pl = &C.lay;
        p_free = C.free_mask;
        p_int = C.free_interior;
        p_tab = std::move(c_tab);
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
    L.make_tiles(coarse_variant(i));
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

// ------------------------------------------------------------------ driver

void Session::smooth(Level& L, int sweeps, int kind, float omega, const float* tau)
{
    // the learned proposal, when set, runs in the fine level's (unshifted)
    // smoothing sweeps as in plain runs; the other Jacobi buffer holds each
    // DOF's state before the last sweep (or correction), its displacement input
    const bool propose = proposal != nullptr && &L == &fine && tau == nullptr && kind != 1;
    for (int s = 0; s < sweeps; ++s) {
        float* xin = L.x[L.cur].p;
        float* xout = L.x[1 - L.cur].p;
        SweepArgs a = L.sweep_args(xin, xout, omega, 0.0f, tau);
        a.counters = counters_for(L);
        if (propose) {
            a.mlp = proposal->view();
            a.prop_mode = proposal_mode;
            a.prop_scale = proposal_scale;
        }
        launch_sweep(dim, kind, L.variant, a, stream);
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
    launch_restrict(dim, t, x_parent, child.x[0].p, parent.grad.p, child.tau.p, child.d_free.p, stream);
    child.cur = 0;
    halo(child, child.x[0].p);
    launch_copy2(child.rx.p, child.x[1].p, child.x[0].p, (long long)child.plane * dim, stream);
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
    const int nv = num_tile_variants(dim);
    std::vector<double> ms(n_coarse * nv, 0.0);
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
                L.make_tiles(v);
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
                if (ms[i * nv + v] < ms[i * nv + best]) best = v;
            picks[i] = best;
            L.make_tiles(best);
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
        for (size_t i = 0; i < n_coarse; ++i) coarse[i]->make_tiles(coarse_variant(i));
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
    // fine correction P(x_c - rx_c), zero where not free; its replicas
    launch_prolong(dim, transfer(F, C0, true), C0.x[C0.cur].p, C0.rx.p, F.corr.p, F.d_free.p, false, nullptr,
                   nullptr, stream);
    halo(F, F.corr.p);
}

void Session::fas_device_ladder(int kind)
{
    Level& F = fine;
    if (!F.structured_energy) throw Internal("run_fas: the line search needs the structured cell energy stencil");
    float* x = F.x[F.cur].p;
    launch_linesearch_ladder(dim, kind, fine_cells(x), F.corr.p, d_ladder_parts.p, (double)1e-6f, d_ladder_out.p,
                             d_alpha.p, stream, F.es_dir.n ? &F.es_dir : nullptr);
    launch_masked_axpy_dev(x, F.corr.p, d_alpha.p, F.d_free.p, F.lay.n_slots(), F.plane, dim, stream);
    halo(F, x);  // a no-op copy when nothing moved (the replicas already equal their owners)
}

void Session::fas_tail(int kind, const FasParams& o, double* out_e, double* out_m)
{
    Level& F = fine;
    smooth(F, o.nu_post, kind, o.omega, nullptr);
    reduce(F, F.x[F.cur].p, kind, 0.0, out_e, out_m);
}

hipGraphExec_t Session::fas_graph(int parity, int kind, const FasParams& o, size_t n_coarse)
{
    FasGraph& g = fas_graphs[parity];
    const bool propose = proposal != nullptr && kind != 1;
    const void* mlp = propose ? (const void*)proposal->view().fwd : nullptr;
    if (g.exec && g.kind == kind && g.variant == tile_variant && g.nu_pre == o.nu_pre && g.nu_post == o.nu_post &&
        g.nu_coarse == o.nu_coarse && g.n_coarse == n_coarse && g.omega == o.omega && g.rep == d_fas_reports.p &&
        g.mlp == mlp && (!propose || (g.mode == proposal_mode && g.scale == proposal_scale)))
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
        fas_tail(kind, o, d_cyc_rep.p, d_cyc_rep.p + 1);
        launch_append_report(d_cyc_rep.p, d_fas_reports.p, (int)(d_fas_reports.n / 2), d_rep_count.p, stream);
    } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        F.cur = cur0;
        for (size_t i = 0; i < coarse.size(); ++i) coarse[i]->cur = ccur[i];
        last_launches = launches0;
        throw;
    }
    EGGMI_HIP(hipStreamEndCapture(stream, &graph));
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    g.launches = last_launches - launches0;
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
    return g.exec;
}

}  // namespace eggmi
