// Host driver of the control-net reduced Gauss-Newton smoother (reference
// src/control_solve.hpp:352-509 set-up, :533-676 run_plain, :736-818 matvec
// and PCG). The loop is the reference's: reduced gradient, Jacobi-
// preconditioned PCG on the exact reduced Hessian plus lam I, a two-trials-
// per-pass line search of M dC on the global fine energy, the damping ladder.
// The PCG's scalars are read back once per reduction (the outer loop's
// decisions need them on the host anyway; the nets are small).
#include "session.hpp"
#include "tmop_math.hpp"

#include <algorithm>
#include <cmath>

namespace eggmi
{

void upload_net_axes(int dim, const CtrlGeom& g, const int32_t* const* off_in, const double* const* wt_in,
                     DArr<int>* d_off, DArr<double>* d_wt, DArr<int>* d_soff, DArr<int>* d_snode, DArr<double>* d_sw)
{
    for (int k = 0; k < dim; ++k) {
        if (!off_in[k] || !wt_in[k]) throw InvalidArgument("control: off / wt of every axis are required");
        std::vector<int> off(off_in[k], off_in[k] + g.n[k]);
        std::vector<double> wt(wt_in[k], wt_in[k] + (size_t)g.n[k] * g.ks);
        std::vector<std::vector<std::pair<int, double>>> lists(g.c[k]);
        for (int i = 0; i < g.n[k]; ++i) {
            if (off[i] < 0 || off[i] + g.ks > g.c[k])
                throw InvalidArgument("control: axis " + std::to_string(k) + " window of node " + std::to_string(i) +
                                      " leaves the lattice");
            for (int t = 0; t < g.ks; ++t) {
                const double w = wt[(size_t)i * g.ks + t];
                if (!std::isfinite(w)) throw InvalidArgument("control: non-finite basis weight");
                if (w != 0.0) lists[off[i] + t].emplace_back(i, w);
            }
        }
        std::vector<int> soff(g.c[k] + 1, 0), snode;
        std::vector<double> sw;
        for (int a = 0; a < g.c[k]; ++a) {
            for (auto& [i, w] : lists[a]) {
                snode.push_back(i);
                sw.push_back(w);
            }
            soff[a + 1] = (int)snode.size();
        }
        d_off[k].upload(off);
        d_wt[k].upload(wt);
        d_soff[k].upload(soff);
        d_snode[k].upload(snode);
        d_sw[k].upload(sw);
    }
}

void Session::set_control(const eggmi_control& c)
{
    need_device();
    const int b = c.block;
    if (b < 0 || b >= (int)fine.lay.shape.size())
        throw InvalidArgument("control: block " + std::to_string(b) + " out of range");
    if (c.k_supp < 1) throw InvalidArgument("control: k_supp must be >= 1");
    auto net = std::make_unique<ControlNet>();
    CtrlGeom& g = net->g;
    g.dim = dim;
    g.ks = c.k_supp;
    g.n_nodes = g.n_ctrl = g.n_cells = 1;
    for (int k = 0; k < dim; ++k) {
        g.n[k] = fine.lay.shape[b][k];
        g.c[k] = c.ctrl_shape[k];
        if (g.c[k] < g.ks) throw InvalidArgument("control: an axis has fewer controls than the support width");
        if (g.n[k] < 2) throw InvalidArgument("control: the block needs at least 2 nodes per axis");
        g.n_nodes *= g.n[k];
        g.n_ctrl *= g.c[k];
        g.n_cells *= g.n[k] - 1;
    }
    if (!c.C0) throw InvalidArgument("control: C0 is required");
    if (c.n_free < 0 || (c.n_free > 0 && !c.free_idx)) throw InvalidArgument("control: free_idx is required");
    g.n_free = (int)c.n_free;
    g.bd = fine.block_desc((size_t)b);
    g.plane = fine.plane;
    net->block = b;
    upload_net_axes(dim, g, c.off, c.wt, net->off, net->wt, net->sup_off, net->sup_node, net->sup_w);
    std::vector<int> fidx(c.free_idx, c.free_idx + g.n_free), l2f(g.n_ctrl, -1);
    for (int f = 0; f < g.n_free; ++f) {
        if (fidx[f] < 0 || fidx[f] >= g.n_ctrl) throw InvalidArgument("control: free_idx out of range");
        if (l2f[fidx[f]] >= 0) throw InvalidArgument("control: free_idx lists a control twice");
        l2f[fidx[f]] = f;
    }
    net->free_idx.upload(fidx);
    net->l2f.upload(l2f);
    net->C.upload(std::vector<double>(c.C0, c.C0 + g.n_ctrl * dim));
    if (c.b) net->b.upload(std::vector<double>(c.b, c.b + g.n_nodes * dim));
    for (int k = 0; k < dim; ++k) {
        g.off[k] = net->off[k].p;
        g.wt[k] = net->wt[k].p;
        g.sup_off[k] = net->sup_off[k].p;
        g.sup_node[k] = net->sup_node[k].p;
        g.sup_w[k] = net->sup_w[k].p;
    }
    g.free_idx = net->free_idx.p;
    g.l2f = net->l2f.p;
    // work buffers
    const long long V = dim * dim, NH = V * (V + 1) / 2, ns = g.n_cells * (1LL << dim);
    const size_t nfd = (size_t)g.n_free * dim;
    net->gs.alloc(ns * V);
    net->zs.alloc(ns * V);
    net->hs.alloc(ns * NH);
    net->xb.alloc(g.n_nodes * dim);
    net->ynode.alloc(g.n_nodes * dim);
    net->dir.alloc(g.n_nodes * dim);
    net->hnode.alloc(g.n_nodes * V);
    net->prec.alloc((size_t)g.n_free * V);
    for (auto* a : {&net->G, &net->dC, &net->r, &net->z, &net->p, &net->Ap}) a->alloc(std::max<size_t>(nfd, 1));
    net->scal.alloc(8);
    net->eparts.alloc(4 * kCtrlParts);
    // the cells of every other block, reduced once per run (their nodes
    // never move: the free controls reach the block's interior nodes only)
    {
        std::vector<BlockDesc> ob;
        std::vector<long long> off {0};
        for (size_t q = 0; q < fine.lay.shape.size(); ++q) {
            if ((int)q == b) continue;
            ob.push_back(fine.block_desc(q));
            long long nc = 1;
            for (int k = 0; k < dim; ++k) nc *= std::max(0, fine.lay.shape[q][k] - 1);
            off.push_back(off.back() + nc);
        }
        net->n_oblocks = (int)ob.size();
        net->n_ocells = off.back();
        if (!ob.empty()) {
            net->oblocks.upload(ob);
            net->ocell_off.upload(off);
        }
    }
    control_topo.reset();
    control = std::move(net);
    control_eval_to_x();
}

void Session::control_eval_to_x()
{
    ControlNet& c = *control;
    float* x = fine.x[fine.cur].p;
    launch_ctrl_eval_state(c.g, c.C.p, c.b.n ? c.b.p : nullptr, c.xb.p, x, stream);
    halo(fine, x);
    pair_valid = false;  // the other Jacobi buffer is stale now
    EGGMI_HIP(hipStreamSynchronize(stream));
}

void Session::control_set_C(const double* C, long long n)
{
    if (control_topo) {
        const long long want = control_topo->n_stack * dim;
        if (n != want)
            throw InvalidArgument("set_C: expected " + std::to_string(want) + " values, got " + std::to_string(n));
        for (long long e = 0; e < n; ++e)
            if (!std::isfinite(C[e])) throw InvalidArgument("set_C: C is not finite");
        EGGMI_HIP(hipMemcpy(control_topo->C.p, C, n * sizeof(double), hipMemcpyHostToDevice));
        control_topo_eval_to_x();
        return;
    }
    if (!control) throw InvalidArgument("set_C: the session has no control wire");
    if (n != control->g.n_ctrl * dim)
        throw InvalidArgument("set_C: expected " + std::to_string(control->g.n_ctrl * dim) + " values, got " +
                              std::to_string(n));
    EGGMI_HIP(hipMemcpy(control->C.p, C, n * sizeof(double), hipMemcpyHostToDevice));
    control_eval_to_x();
}

void Session::control_get_C(double* C, long long n) const
{
    if (control_topo) {
        const long long want = control_topo->n_stack * dim;
        if (n != want)
            throw InvalidArgument("get_C: expected " + std::to_string(want) + " values, got " + std::to_string(n));
        EGGMI_HIP(hipMemcpy(C, control_topo->C.p, n * sizeof(double), hipMemcpyDeviceToHost));
        return;
    }
    if (!control) throw InvalidArgument("get_C: the session has no control wire");
    if (n != control->g.n_ctrl * dim)
        throw InvalidArgument("get_C: expected " + std::to_string(control->g.n_ctrl * dim) + " values, got " +
                              std::to_string(n));
    EGGMI_HIP(hipMemcpy(C, control->C.p, n * sizeof(double), hipMemcpyDeviceToHost));
}

ControlReport Session::run_control(const ControlOptions& o, int kind)
{
    need_device();
    if (control_topo) return run_control_topo(o, kind);
    if (!control) throw InvalidArgument("run_control: the session has no control wire");
    if (kind == kUntangle) throw InvalidArgument("run_control: the untangle phase stays nodal (barrier phases only)");
    if (!fine.structured_energy)
        throw InvalidArgument("run_control: the line search needs the structured cell energy stencil");
    if (weighted) throw InvalidArgument("run_control: per-sample energy weights are not supported");
    if (fine.es_dir.n) throw InvalidArgument("run_control: directional energy terms are not supported");
    ControlNet& c = *control;
    const CtrlGeom& g = c.g;
    const long long nfd = (long long)g.n_free * dim;
    const double tol = 1e-6;  // FP32 accept slack (reference src/core.hpp:40)
    ControlReport rep;
    double s[4];
    auto scalar = [&](int i) {
        double v = 0.0;
        EGGMI_HIP(hipMemcpyAsync(&v, c.scal.p + i, sizeof(double), hipMemcpyDeviceToHost, stream));
        EGGMI_HIP(hipStreamSynchronize(stream));
        return v;
    };
    // (A + lam I) v -> out: direction field M v, the per-sample Hessian
    // action gathered per node, pulled back to the free controls; inside
    // the PCG, so the kernels return at once once it has stopped
    CtrlGeom gh = g;
    gh.halt = c.scal.p + 6;
    auto matvec = [&](double lam, const double* v, double* out) {
        launch_ctrl_eval_dir(gh, v, c.dir.p, nullptr, stream);
        launch_ctrl_cell(gh, true, nullptr, c.hs.p, c.dir.p, c.zs.p, stream);
        launch_ctrl_node_cc(gh, c.zs.p, c.ynode.p, stream);
        launch_ctrl_pullback(gh, 0, c.ynode.p, out, stream);
        if (lam > 0.0) launch_ctrl_axpy(nfd, lam, v, out, false, stream);
    };
    // the reference's Jacobi-preconditioned CG (src/control_solve.hpp:
    // 761-818) with its scalars on the device: chunks of iterations, one
    // 8-double read-back each, the chunk sized from the residual's observed
    // rate to the stop threshold
    double ps[8];
    auto read_state = [&]() {
        EGGMI_HIP(hipMemcpyAsync(ps, c.scal.p, sizeof(ps), hipMemcpyDeviceToHost, stream));
        EGGMI_HIP(hipStreamSynchronize(stream));
    };
    auto pcg = [&](double lam, int max_iter, double rtol) -> std::pair<int, double> {
        launch_qpcg_init(dim, true, nfd, c.G.p, c.prec.p, lam, rtol, c.dC.p, c.r.p, c.z.p, c.p.p, c.scal.p, stream);
        read_state();
        if (ps[3] == 0.0) return {0, 0.0};
        if (max_iter <= 0) return {0, 1.0};
        double rr_prev = ps[3];
        int chunk = 4, done_it = 0;
        while (true) {
            const int n = std::max(1, std::min(chunk, max_iter - done_it));
            for (int i = 0; i < n; ++i) {
                matvec(lam, c.p.p, c.Ap.p);
                launch_qpcg_pap(nfd, c.p.p, c.Ap.p, c.scal.p, stream);
                launch_qpcg_update(dim, true, nfd, c.prec.p, lam, max_iter, c.dC.p, c.r.p, c.z.p, c.p.p, c.Ap.p,
                                   c.scal.p, stream);
            }
            read_state();
            const int it = (int)ps[5];
            if (ps[6] != 0.0 || it >= max_iter) break;
            const double rate = std::pow(ps[1] / rr_prev, 1.0 / std::max(1, it - done_it));
            chunk = 8;
            if (rate > 0.0 && rate < 1.0 && ps[1] > ps[4])
                chunk = (int)std::clamp(std::ceil(std::log(ps[4] / ps[1]) / std::log(rate)), 2.0, 32.0);
            rr_prev = ps[1];
            done_it = it;
        }
        return {(int)ps[5], std::sqrt(ps[1] / ps[3])};
    };

    // the net owns the block from here: start exactly on its manifold
    control_eval_to_x();
    // the global energy: the other blocks' cells (constant) plus the block's,
    // which the loop evaluates in FP64 from the net's positions
    double e_other = 0.0, md_other = HUGE_VAL;
    if (c.n_ocells > 0) {
        CellReduceArgs ra = fine_cells(fine.x[fine.cur].p);
        ra.blocks = c.oblocks.p;
        ra.cell_off = c.ocell_off.p;
        ra.nblocks = c.n_oblocks;
        ra.ncells = c.n_ocells;
        launch_reduce_cells(dim, kind, ra, partials.p, stream);
        launch_reduce_finish(partials.p, kReduceBlocks, c.scal.p, c.scal.p + 1, stream);
        e_other = scalar(0);
        md_other = scalar(1);
    }
    launch_ctrl_energy(g, kind, c.xb.p, nullptr, 0.0, 0.0, c.eparts.p, c.scal.p, stream);
    double e0 = e_other + scalar(0), md0 = std::min(md_other, scalar(1));
    rep.energies.push_back(e0);
    rep.mindets.push_back(md0);
    double lam = o.lm0, gmax0 = 0.0;
    for (int outer = 0; outer < o.n_outer; ++outer) {
        // reduced gradient G = M_f^T g from the per-sample derivatives
        launch_ctrl_samples(g, kind, c.xb.p, c.gs.p, c.hs.p, stream);
        launch_ctrl_cell(g, false, c.gs.p, nullptr, nullptr, c.zs.p, stream);
        launch_ctrl_node_cc(g, c.zs.p, c.ynode.p, stream);
        launch_ctrl_pullback(g, 0, c.ynode.p, c.G.p, stream);
        launch_ctrl_reduce(1, c.G.p, nullptr, nfd, dim, c.scal.p, stream);
        const double gmax = scalar(0);
        if (gmax < o.grad_tol) {
            rep.converged = true;
            break;
        }
        if (outer == 0) gmax0 = gmax;
        const double pcg_tol =
            o.pcg_forcing ? std::max(o.pcg_rtol, std::min(0.5, std::sqrt(gmax / gmax0))) : o.pcg_rtol;
        // Jacobi preconditioner from the contracted node Hessians; damping floor
        launch_ctrl_node_hessians(g, c.hs.p, c.zs.p, c.hnode.p, stream);
        launch_ctrl_pullback(g, 1, c.hnode.p, c.prec.p, stream);
        launch_ctrl_reduce(2, c.prec.p, nullptr, g.n_free, dim, c.scal.p, stream);
        const double lam_floor = 1e-8 * scalar(0) / (double)std::max<long long>(nfd, 1);

        bool accepted = false;
        double alpha = 1.0, e_new = e0, md_new = md0;
        for (int lt = 0; lt < o.lm_tries && !accepted; ++lt) {
            const auto [its, relres] = pcg(lam, o.pcg_max_iter, pcg_tol);
            rep.pcg_iters.push_back((double)its);
            rep.pcg_relres.push_back(relres);
            // the node-space step M dC, line-searched two trials per pass
            launch_ctrl_eval_dir(g, c.dC.p, c.dir.p, nullptr, stream);
            alpha = 1.0;
            while (alpha >= o.alpha_min) {
                launch_ctrl_energy(g, kind, c.xb.p, c.dir.p, alpha, alpha * 0.5, c.eparts.p, c.scal.p, stream);
                EGGMI_HIP(hipMemcpyAsync(s, c.scal.p, sizeof(s), hipMemcpyDeviceToHost, stream));
                EGGMI_HIP(hipStreamSynchronize(stream));
                bool ok = false;
                for (int t = 0; t < 2 && !ok; ++t) {
                    const double e = e_other + s[2 * t], md = std::min(md_other, s[2 * t + 1]);
                    const double a = t == 0 ? alpha : alpha * 0.5;
                    if (a >= o.alpha_min && std::isfinite(e) && e <= e0 + tol && md > 0.0) {
                        alpha = a;
                        e_new = e;
                        md_new = md;
                        ok = true;
                    }
                }
                if (ok) {
                    accepted = true;
                    break;
                }
                alpha *= 0.25;
            }
            if (!accepted) lam = std::max(10.0 * lam, lam_floor);
        }
        if (!accepted) break;  // damping exhausted
        // commit: C_free += alpha dC, the net re-evaluated into the FP64
        // positions and the store (with the block's replicas)
        launch_ctrl_scatter_free(g, alpha, c.dC.p, c.C.p, stream);
        launch_ctrl_eval_state(g, c.C.p, c.b.n ? c.b.p : nullptr, c.xb.p, fine.x[fine.cur].p, stream);
        halo(fine, fine.x[fine.cur].p);
        e0 = e_new;
        md0 = md_new;
        rep.energies.push_back(e0);
        rep.mindets.push_back(md0);
        rep.alphas.push_back(alpha);
        rep.iters += 1;
        if (alpha == 1.0) lam /= 3.0;
    }
    EGGMI_HIP(hipStreamSynchronize(stream));
    pair_valid = false;
    return rep;
}

}  // namespace eggmi
