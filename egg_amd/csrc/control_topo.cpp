// Host driver of the control-net reduced Gauss-Newton smoother over several
// blocks' nets (reference ControlTopoSessionT, src/control_solve.hpp:
// 1585-2837: set-up :1590-2055, set_reduction / set_penalty / set_b
// :2057-2184, run :2209-2425, matvec_q and pcg_solve_q :2529-2606). The
// nets share one stacked control vector C whose step is dC = R dq over the
// reduced DOFs q; homogeneous penalty rows 1/2 sum w (P C)^2 join the
// Gauss-Newton system and the line-search energy. Every block of the session
// is a domain of cells: the positions live in FP64 over the concatenated
// block nodes, a node shared by several blocks takes its owner's value and
// the gathers credit the owner with its copies' contributions.
#include "session.hpp"
#include "tmop_math.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>

namespace eggmi
{

namespace
{

// CSR checks shared by the reduction and the penalty: off has rows + 1
// non-decreasing entries from 0, every column below cols, coefficients finite
void check_csr(const char* what, long long rows, long long cols, const int* off, const int* col, const double* coef)
{
    if (rows > 0 && !off) throw InvalidArgument(std::string("control topo: ") + what + " offsets are required");
    if (rows == 0) return;
    if (off[0] != 0) throw InvalidArgument(std::string("control topo: ") + what + " offsets must start at 0");
    for (long long i = 0; i < rows; ++i)
        if (off[i + 1] < off[i]) throw InvalidArgument(std::string("control topo: ") + what + " offsets decrease");
    const long long ne = off[rows];
    if (ne > 0 && (!col || !coef))
        throw InvalidArgument(std::string("control topo: ") + what + " columns and coefficients are required");
    for (long long e = 0; e < ne; ++e) {
        if (col[e] < 0 || col[e] >= cols)
            throw InvalidArgument(std::string("control topo: ") + what + " column out of range");
        if (!std::isfinite(coef[e]))
            throw InvalidArgument(std::string("control topo: ") + what + " coefficient not finite");
    }
}

// transpose of a rows x cols CSR: per column its (row, coef) entries, rows
// ascending
void transpose_csr(long long rows, long long cols, const int* off, const int* col, const double* coef,
                   std::vector<int>& t_off, std::vector<int>& t_row, std::vector<double>& t_coef)
{
    t_off.assign(cols + 1, 0);
    const long long ne = rows > 0 ? off[rows] : 0;
    for (long long e = 0; e < ne; ++e) ++t_off[col[e] + 1];
    for (long long c = 0; c < cols; ++c) t_off[c + 1] += t_off[c];
    t_row.assign(ne, 0);
    t_coef.assign(ne, 0.0);
    std::vector<int> cur(t_off.begin(), t_off.end() - 1);
    for (long long r = 0; r < rows; ++r)
        for (int e = off[r]; e < off[r + 1]; ++e) {
            const int k = cur[col[e]]++;
            t_row[k] = (int)r;
            t_coef[k] = coef[e];
        }
}

}  // namespace

void Session::set_control_topo(const eggmi_control_topo& t)
{
    need_device();
    const int nb = (int)fine.lay.shape.size();
    if (t.n_nets < 1 || !t.nets) throw InvalidArgument("control topo: at least one net is required");
    if (!t.C0) throw InvalidArgument("control topo: C0 is required");
    auto T = std::make_unique<ControlTopo>();
    ControlTopo& c = *T;
    c.n_nets = t.n_nets;
    // every block a domain of cells, in block order
    c.hblk.resize(nb);
    std::vector<long long> npfx {0}, spfx {0};
    for (int b = 0; b < nb; ++b) {
        TopoBlock& B = c.hblk[b];
        CtrlGeom& g = B.g;
        g.dim = dim;
        g.n_nodes = g.n_cells = 1;
        for (int k = 0; k < dim; ++k) {
            g.n[k] = fine.lay.shape[b][k];
            g.n_nodes *= g.n[k];
            g.n_cells *= std::max(0, g.n[k] - 1);
        }
        g.bd = fine.block_desc((size_t)b);
        g.plane = fine.plane;
        B.node0 = npfx.back();
        B.samp0 = spfx.back();
        npfx.push_back(B.node0 + g.n_nodes);
        spfx.push_back(B.samp0 + g.n_cells * (1LL << dim));
    }
    c.n_all = npfx.back();
    c.ns_all = spfx.back();
    if (c.n_all >= INT_MAX) throw InvalidArgument("control topo: too many block nodes");
    // the nets: all their controls free (the reduction carries the freedom)
    c.off.resize(c.n_nets);
    c.wt.resize(c.n_nets);
    c.sup_off.resize(c.n_nets);
    c.sup_node.resize(c.n_nets);
    c.sup_w.resize(c.n_nets);
    c.b.resize(c.n_nets);
    std::vector<long long> cpfx {0};
    for (int i = 0; i < c.n_nets; ++i) {
        const eggmi_control& w = t.nets[i];
        if (w.block < 0 || w.block >= nb)
            throw InvalidArgument("control topo: net " + std::to_string(i) + " names block " +
                                  std::to_string(w.block) + ", out of range");
        TopoBlock& B = c.hblk[w.block];
        if (B.net >= 0) throw InvalidArgument("control topo: block " + std::to_string(w.block) + " has two nets");
        if (w.k_supp < 1) throw InvalidArgument("control topo: k_supp must be >= 1");
        CtrlGeom& g = B.g;
        g.ks = w.k_supp;
        g.n_ctrl = 1;
        for (int k = 0; k < dim; ++k) {
            g.c[k] = w.ctrl_shape[k];
            if (g.c[k] < 2 || g.c[k] < g.ks)
                throw InvalidArgument("control topo: each ctrl_shape entry must be >= 2 and >= k_supp");
            if (g.n[k] < 2) throw InvalidArgument("control topo: a net's block needs at least 2 nodes per axis");
            g.n_ctrl *= g.c[k];
        }
        g.n_free = (int)g.n_ctrl;
        upload_net_axes(dim, g, w.off, w.wt, c.off[i].data(), c.wt[i].data(), c.sup_off[i].data(),
                        c.sup_node[i].data(), c.sup_w[i].data());
        for (int k = 0; k < dim; ++k) {
            g.off[k] = c.off[i][k].p;
            g.wt[k] = c.wt[i][k].p;
            g.sup_off[k] = c.sup_off[i][k].p;
            g.sup_node[k] = c.sup_node[i][k].p;
            g.sup_w[k] = c.sup_w[i][k].p;
        }
        if (w.b) {
            for (long long e = 0; e < g.n_nodes * dim; ++e)
                if (!std::isfinite(w.b[e])) throw InvalidArgument("control topo: b is not finite");
            c.b[i].upload(std::vector<double>(w.b, w.b + g.n_nodes * dim));
        }
        B.b = c.b[i].p;
        B.net = i;
        B.ctrl0 = cpfx.back();
        cpfx.push_back(B.ctrl0 + g.n_ctrl);
        c.net_block.push_back(w.block);
        c.net_nodes.push_back(g.n_nodes);
    }
    c.n_stack = cpfx.back();
    const long long nsd = c.n_stack * dim;
    if (nsd >= INT_MAX) throw InvalidArgument("control topo: too many controls");
    std::vector<double> C0(t.C0, t.C0 + nsd);
    for (double v : C0)
        if (!std::isfinite(v)) throw InvalidArgument("control topo: C0 is not finite");
    // owners: the block node at its global point's root slot; the others
    // are its copies
    std::vector<int> owner(c.n_all, -1);
    {
        std::vector<int> root_node(M, -1);
        for (int b = 0; b < nb; ++b) {
            const CtrlGeom& g = c.hblk[b].g;
            for (long long n = 0; n < g.n_nodes; ++n) {
                int I[3] = {0, 0, 0};
                long long r = n;
                for (int k = dim - 1; k >= 0; --k) {
                    I[k] = int(r % g.n[k]);
                    r /= g.n[k];
                }
                const int gid = bgd[b][n];
                if (fine.lay.interior_slot(b, I) == g2s[gid]) root_node[gid] = (int)(c.hblk[b].node0 + n);
            }
        }
        for (int b = 0; b < nb; ++b)
            for (long long n = 0; n < c.hblk[b].g.n_nodes; ++n) {
                const int gid = bgd[b][n];
                if (root_node[gid] < 0)
                    throw Internal("control topo: global node " + std::to_string(gid) +
                                   " has no block node at its root slot");
                owner[c.hblk[b].node0 + n] = root_node[gid];
            }
    }
    std::vector<int> grp_off(c.n_all + 1, 0), grp_mem;
    for (long long n = 0; n < c.n_all; ++n)
        if (owner[n] != n) ++grp_off[owner[n] + 1];
    for (long long n = 0; n < c.n_all; ++n) grp_off[n + 1] += grp_off[n];
    grp_mem.assign(grp_off[c.n_all], 0);
    {
        std::vector<int> cur(grp_off.begin(), grp_off.end() - 1);
        for (long long n = 0; n < c.n_all; ++n)
            if (owner[n] != n) grp_mem[cur[owner[n]]++] = (int)n;  // ascending per owner
    }
    // device tables and the view
    c.blk.upload(c.hblk);
    c.npfx.upload(npfx);
    c.spfx.upload(spfx);
    c.cpfx.upload(cpfx);
    c.net_blk.upload(c.net_block);
    c.owner.upload(owner);
    c.grp_off.upload(grp_off);
    c.grp_mem.upload(grp_mem.empty() ? std::vector<int> {0} : grp_mem);
    TopoView& v = c.view;
    v.blk = c.blk.p;
    v.nb = nb;
    v.npfx = c.npfx.p;
    v.spfx = c.spfx.p;
    v.n_nets = c.n_nets;
    v.net_blk = c.net_blk.p;
    v.cpfx = c.cpfx.p;
    v.owner = c.owner.p;
    v.grp_off = c.grp_off.p;
    v.grp_mem = c.grp_mem.p;
    v.n_all = c.n_all;
    v.ns_all = c.ns_all;
    v.n_stack = c.n_stack;
    v.dim = dim;
    // work buffers
    const long long V = dim * dim, NH = V * (V + 1) / 2;
    c.C.upload(C0);
    for (auto* a : {&c.xb, &c.dir, &c.ynode}) a->alloc(std::max<long long>(c.n_all * dim, 1));
    c.hnode.alloc(std::max<long long>(c.n_all * V, 1));
    c.gs.alloc(std::max<long long>(c.ns_all * V, 1));
    c.zs.alloc(std::max<long long>(c.ns_all * V, 1));
    c.hs.alloc(std::max<long long>(c.ns_all * NH, 1));
    c.prec.alloc(std::max<long long>(c.n_stack * V, 1));
    for (auto* a : {&c.vfull, &c.wfull, &c.dCfull}) a->alloc(std::max<long long>(nsd, 1));
    c.scal.alloc(8);
    c.eparts.alloc(4 * kCtrlParts);
    control.reset();
    control_topo = std::move(T);
    try {
        control_topo_set_reduction(t.nq, t.r_off, t.r_col, t.r_coef);
        control_topo_set_penalty(t.n_pen, t.p_off, t.p_col, t.p_coef, t.p_w);
    } catch (...) {
        control_topo.reset();
        throw;
    }
    control_topo_eval_to_x();
}

void Session::control_topo_set_reduction(long long nq, const int* off, const int* col, const double* coef)
{
    if (!control_topo) throw InvalidArgument("set_control_reduction: the session has no multi-block control wire");
    ControlTopo& c = *control_topo;
    const long long nsd = c.n_stack * dim;
    if (nq < 0 || nq >= INT_MAX) throw InvalidArgument("control topo: nq out of range");
    check_csr("reduction", nsd, nq, off, col, coef);
    const long long ne = off[nsd];
    std::vector<int> o(off, off + nsd + 1), cl(col, col + ne), g_off, g_row;
    std::vector<double> cf(coef, coef + ne), g_coef;
    transpose_csr(nsd, nq, off, col, coef, g_off, g_row, g_coef);
    c.nq = nq;
    c.r_off.upload(o);
    c.r_col.upload(cl.empty() ? std::vector<int> {0} : cl);
    c.r_coef.upload(cf.empty() ? std::vector<double> {0.0} : cf);
    c.g_off.upload(g_off);
    c.g_row.upload(g_row.empty() ? std::vector<int> {0} : g_row);
    c.g_coef.upload(g_coef.empty() ? std::vector<double> {0.0} : g_coef);
    for (auto* a : {&c.qG, &c.qdq, &c.qr, &c.qz, &c.qp, &c.qAp, &c.qdiag}) a->alloc(std::max<long long>(nq, 1));
}

void Session::control_topo_set_penalty(long long n_pen, const int* off, const int* col, const double* coef,
                                       const double* w)
{
    if (!control_topo) throw InvalidArgument("set_control_penalty: the session has no multi-block control wire");
    ControlTopo& c = *control_topo;
    const long long nsd = c.n_stack * dim;
    if (n_pen < 0 || n_pen >= INT_MAX) throw InvalidArgument("control topo: penalty row count out of range");
    if (n_pen == 0) {
        c.n_pen = 0;
        c.p_w_h.clear();
        return;
    }
    check_csr("penalty", n_pen, nsd, off, col, coef);
    if (!w) throw InvalidArgument("control topo: penalty weights are required");
    for (long long j = 0; j < n_pen; ++j)
        if (!std::isfinite(w[j])) throw InvalidArgument("control topo: penalty weight not finite");
    const long long ne = off[n_pen];
    std::vector<int> pt_off, pt_row;
    std::vector<double> pt_coef;
    transpose_csr(n_pen, nsd, off, col, coef, pt_off, pt_row, pt_coef);
    c.p_off.upload(std::vector<int>(off, off + n_pen + 1));
    c.p_col.upload(ne ? std::vector<int>(col, col + ne) : std::vector<int> {0});
    c.p_coef.upload(ne ? std::vector<double>(coef, coef + ne) : std::vector<double> {0.0});
    c.p_w_h.assign(w, w + n_pen);
    c.p_w.upload(c.p_w_h);
    c.pt_off.upload(pt_off);
    c.pt_row.upload(pt_row.empty() ? std::vector<int> {0} : pt_row);
    c.pt_coef.upload(pt_coef.empty() ? std::vector<double> {0.0} : pt_coef);
    c.r_pen.alloc(n_pen);
    c.n_pen = n_pen;
}

void Session::control_topo_set_b(int net, const double* b, long long n)
{
    if (!control_topo) throw InvalidArgument("set_control_b: the session has no multi-block control wire");
    ControlTopo& c = *control_topo;
    if (net < 0 || net >= c.n_nets) throw InvalidArgument("set_control_b: net index out of range");
    if (!b || n != c.net_nodes[net] * dim)
        throw InvalidArgument("set_control_b: expected " + std::to_string(c.net_nodes[net] * dim) + " values");
    for (long long e = 0; e < n; ++e)
        if (!std::isfinite(b[e])) throw InvalidArgument("set_control_b: b is not finite");
    c.b[net].upload(std::vector<double>(b, b + n));
    // the block's record points at the new array
    TopoBlock& B = c.hblk[c.net_block[net]];
    B.b = c.b[net].p;
    EGGMI_HIP(hipMemcpy(c.blk.p + c.net_block[net], &B, sizeof(TopoBlock), hipMemcpyHostToDevice));
    control_topo_eval_to_x();
}

void Session::control_topo_eval_to_x()
{
    ControlTopo& c = *control_topo;
    float* x = fine.x[fine.cur].p;
    launch_topo_load(c.view, x, c.xb.p, stream);
    launch_topo_eval(c.view, false, c.C.p, c.xb.p, x, stream);
    halo(fine, x);
    pair_valid = false;
    EGGMI_HIP(hipStreamSynchronize(stream));
}

ControlReport Session::run_control_topo(const ControlOptions& o, int kind)
{
    need_device();
    if (kind == kUntangle) throw InvalidArgument("run_control: the untangle phase stays nodal (barrier phases only)");
    if (!fine.structured_energy)
        throw InvalidArgument("run_control: the line search needs the structured cell energy stencil");
    if (weighted) throw InvalidArgument("run_control: per-sample energy weights are not supported");
    if (fine.es_dir.n) throw InvalidArgument("run_control: directional energy terms are not supported");
    ControlTopo& c = *control_topo;
    const TopoView& tv = c.view;
    const long long nsd = c.n_stack * dim, nq = c.nq;
    const double tol = 1e-6;  // FP32 accept slack (reference src/core.hpp:40)
    ControlReport rep;
    double s[4];
    auto scalar = [&](int i) {
        double v = 0.0;
        EGGMI_HIP(hipMemcpyAsync(&v, c.scal.p + i, sizeof(double), hipMemcpyDeviceToHost, stream));
        EGGMI_HIP(hipStreamSynchronize(stream));
        return v;
    };
    // the penalty residuals P x into host memory
    std::vector<double> r0(c.n_pen), r1(c.n_pen);
    auto pen_residual = [&](const double* x, std::vector<double>& out) {
        launch_csr_apply(c.n_pen, c.p_off.p, c.p_col.p, c.p_coef.p, x, 0.0, nullptr, c.r_pen.p, stream);
        EGGMI_HIP(hipMemcpyAsync(out.data(), c.r_pen.p, c.n_pen * sizeof(double), hipMemcpyDeviceToHost, stream));
        EGGMI_HIP(hipStreamSynchronize(stream));
    };
    auto pen_energy = [&](double a) {  // 1/2 sum w (r0 + a r1)^2
        double e = 0.0;
        for (long long j = 0; j < c.n_pen; ++j) {
            const double v = r0[j] + a * r1[j];
            e += 0.5 * c.p_w_h[j] * v * v;
        }
        return e;
    };
    // the penalty gradient / Hessian action P^T W P x added into wfull
    auto pen_accum = [&](const double* x) {
        if (!c.n_pen) return;
        launch_csr_apply(c.n_pen, c.p_off.p, c.p_col.p, c.p_coef.p, x, 0.0, nullptr, c.r_pen.p, stream);
        launch_csrT_accum(nsd, c.pt_off.p, c.pt_row.p, c.pt_coef.p, c.p_w.p, c.r_pen.p, c.wfull.p, stream);
    };
    // (R^T (A + P^T W P) R + lam I) v -> out, inside the PCG: its heavy
    // kernels return at once once the PCG has stopped
    TopoView tvh = tv;
    tvh.halt = c.scal.p + 6;
    auto matvec = [&](double lam, const double* v, double* out) {
        launch_csr_apply(nsd, c.r_off.p, c.r_col.p, c.r_coef.p, v, 0.0, nullptr, c.vfull.p, stream);
        launch_topo_eval(tvh, true, c.vfull.p, c.dir.p, nullptr, stream);
        launch_topo_cell(tvh, true, nullptr, c.hs.p, c.dir.p, c.zs.p, stream);
        launch_topo_node_cc(tvh, c.zs.p, c.ynode.p, stream);
        launch_topo_pullback(tvh, 0, c.ynode.p, c.wfull.p, stream);
        pen_accum(c.vfull.p);
        launch_csr_apply(nq, c.g_off.p, c.g_row.p, c.g_coef.p, c.wfull.p, lam, lam > 0.0 ? v : nullptr, out, stream);
    };
    // PCG with device-resident scalars: the recurrences run in chunks of
    // iterations with one 8-double read-back per chunk; the chunk length
    // follows the residual's observed rate to the stop threshold (the
    // iterations past the stop are no-ops on the device)
    double ps[8];
    auto read_state = [&]() {
        EGGMI_HIP(hipMemcpyAsync(ps, c.scal.p, sizeof(ps), hipMemcpyDeviceToHost, stream));
        EGGMI_HIP(hipStreamSynchronize(stream));
    };
    auto pcg = [&](double lam, int max_iter, double rtol) -> std::pair<int, double> {
        launch_qpcg_init(dim, false, nq, c.qG.p, c.qdiag.p, lam, rtol, c.qdq.p, c.qr.p, c.qz.p, c.qp.p, c.scal.p, stream);
        read_state();
        if (ps[3] == 0.0) return {0, 0.0};
        if (max_iter <= 0) return {0, 1.0};
        double rr_prev = ps[3];
        int chunk = 4, done_it = 0;
        while (true) {
            const int n = std::max(1, std::min(chunk, max_iter - done_it));
            for (int i = 0; i < n; ++i) {
                matvec(lam, c.qp.p, c.qAp.p);
                launch_qpcg_pap(nq, c.qp.p, c.qAp.p, c.scal.p, stream);
                launch_qpcg_update(dim, false, nq, c.qdiag.p, lam, max_iter, c.qdq.p, c.qr.p, c.qz.p, c.qp.p, c.qAp.p, c.scal.p,
                                   stream);
            }
            read_state();
            const int it = (int)ps[5];
            if (ps[6] != 0.0 || it >= max_iter) break;
            // iterations to the threshold at the chunk's mean rate, 2 to 32
            const double rate = std::pow(ps[1] / rr_prev, 1.0 / std::max(1, it - done_it));
            chunk = 8;
            if (rate > 0.0 && rate < 1.0 && ps[1] > ps[4])
                chunk = (int)std::clamp(std::ceil(std::log(ps[4] / ps[1]) / std::log(rate)), 2.0, 32.0);
            rr_prev = ps[1];
            done_it = it;
        }
        return {(int)ps[5], std::sqrt(ps[1] / ps[3])};
    };

    // start exactly on the nets' manifold
    control_topo_eval_to_x();
    launch_topo_energy(tv, kind, c.xb.p, nullptr, 0.0, 0.0, c.eparts.p, c.scal.p, stream);
    double e0 = scalar(0), md0 = scalar(1);
    if (c.n_pen) {
        pen_residual(c.C.p, r0);
        std::fill(r1.begin(), r1.end(), 0.0);
        e0 += pen_energy(0.0);
    }
    rep.energies.push_back(e0);
    rep.mindets.push_back(md0);
    if (nq == 0) return rep;
    double lam = o.lm0, gmax0 = 0.0;
    for (int outer = 0; outer < o.n_outer; ++outer) {
        // reduced gradient qG = R^T (M^T g + P^T W P C)
        launch_topo_samples(tv, kind, c.xb.p, c.gs.p, c.hs.p, stream);
        launch_topo_cell(tv, false, c.gs.p, nullptr, nullptr, c.zs.p, stream);
        launch_topo_node_cc(tv, c.zs.p, c.ynode.p, stream);
        launch_topo_pullback(tv, 0, c.ynode.p, c.wfull.p, stream);
        pen_accum(c.C.p);
        launch_csr_apply(nq, c.g_off.p, c.g_row.p, c.g_coef.p, c.wfull.p, 0.0, nullptr, c.qG.p, stream);
        launch_ctrl_reduce(1, c.qG.p, nullptr, nq, dim, c.scal.p, stream);
        const double gmax = scalar(0);
        if (gmax < o.grad_tol) {
            rep.converged = true;
            break;
        }
        if (outer == 0) gmax0 = gmax;
        const double pcg_tol =
            o.pcg_forcing ? std::max(o.pcg_rtol, std::min(0.5, std::sqrt(gmax / gmax0))) : o.pcg_rtol;
        // per-control Jacobi blocks from the owners' contracted Hessians, the
        // reduced diagonal through R; damping floor from their trace
        launch_topo_node_hessians(tv, c.hs.p, c.zs.p, c.hnode.p, stream);
        launch_topo_pullback(tv, 1, c.hnode.p, c.prec.p, stream);
        launch_qdiag_build(dim, nq, c.g_off.p, c.g_row.p, c.g_coef.p, c.prec.p, c.qdiag.p, stream);
        launch_ctrl_reduce(2, c.prec.p, nullptr, c.n_stack, dim, c.scal.p, stream);
        const double lam_floor = 1e-8 * scalar(0) / (double)std::max<long long>(nsd, 1);

        bool accepted = false;
        double alpha = 1.0, e_new = e0, md_new = md0;
        for (int lt = 0; lt < o.lm_tries && !accepted; ++lt) {
            const auto [its, relres] = pcg(lam, o.pcg_max_iter, pcg_tol);
            rep.pcg_iters.push_back((double)its);
            rep.pcg_relres.push_back(relres);
            // dC = R dq, its node-space step, the penalty residuals along it
            launch_csr_apply(nsd, c.r_off.p, c.r_col.p, c.r_coef.p, c.qdq.p, 0.0, nullptr, c.dCfull.p, stream);
            launch_topo_eval(tv, true, c.dCfull.p, c.dir.p, nullptr, stream);
            if (c.n_pen) {
                pen_residual(c.C.p, r0);
                pen_residual(c.dCfull.p, r1);
            }
            alpha = 1.0;
            while (alpha >= o.alpha_min) {
                launch_topo_energy(tv, kind, c.xb.p, c.dir.p, alpha, alpha * 0.5, c.eparts.p, c.scal.p, stream);
                EGGMI_HIP(hipMemcpyAsync(s, c.scal.p, sizeof(s), hipMemcpyDeviceToHost, stream));
                EGGMI_HIP(hipStreamSynchronize(stream));
                bool ok = false;
                for (int t = 0; t < 2 && !ok; ++t) {
                    const double a = t == 0 ? alpha : alpha * 0.5;
                    const double e = s[2 * t] + (c.n_pen ? pen_energy(a) : 0.0), md = s[2 * t + 1];
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
        // commit: C += alpha dC, the nets re-evaluated into the FP64
        // positions and the store, the store's copies refreshed
        launch_ctrl_axpy(nsd, alpha, c.dCfull.p, c.C.p, false, stream);
        launch_topo_eval(tv, false, c.C.p, c.xb.p, fine.x[fine.cur].p, stream);
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
