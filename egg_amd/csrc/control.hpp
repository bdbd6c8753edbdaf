// Control-net reduced Gauss-Newton smoother on one block (reference
// src/control_solve.hpp:308-342, :533-676, kernels src/control_net.hpp:
// 486-828): the block's nodes are X = M C + b, M the tensor product of
// per-axis banded B-spline windows; each outer iteration pulls the fine
// TMOP gradient back to the free controls (G = M_f^T g), solves
// (A + lambda I) dC = -G with A = sum over samples of W_p^T H_p W_p (the
// exact Hessian of E(C): the map is linear) by Jacobi-preconditioned
// matrix-free PCG in FP64, and line-searches the node-space step M dC on the
// global fine energy under the sweep's accept rule; a rejected step raises
// the Levenberg damping.
//
// Device data: per-axis windows (off, wt) and their inverse (control ->
// nodes with weights), the lattice C and offset b in FP64, the block's node
// positions M C + b in FP64 (the loop's state: an FP32 store would floor
// the reduced gradient near 1e-6, far above the loop's 1e-8 convergence
// test; the session's FP32 store receives the rounded positions after each
// accepted step), per-sample metric gradient and full symmetric Hessian
// (rebuilt once per outer iteration), block-local FP64 node fields for the
// gradient, the step M dC, the Hessian action and the contracted node
// Hessians, and FP64 free-control vectors.
#pragma once

#include "kernels.hpp"

#include <cstdint>
#include <vector>

namespace eggmi
{

// Block and map geometry the kernels share (by value).
struct CtrlGeom {
    int dim = 2;
    int n[3] = {1, 1, 1};   // node shape (interior nodes of the block)
    int c[3] = {1, 1, 1};   // control lattice shape
    int ks = 4;             // support window width per axis
    long long n_nodes = 0;  // prod(n)
    long long n_ctrl = 0;   // prod(c)
    long long n_cells = 0;  // prod(n - 1)
    int n_free = 0;
    BlockDesc bd {};        // the block in the fine store
    long long plane = 0;
    const int* off[3] = {};       // [n[k]] first control of node i's window
    const double* wt[3] = {};     // [n[k] * ks]
    const int* sup_off[3] = {};   // [c[k] + 1] control a -> its nodes
    const int* sup_node[3] = {};
    const double* sup_w[3] = {};
    const int* free_idx = nullptr;  // [n_free] lattice index
    const int* l2f = nullptr;       // [n_ctrl] lattice -> free slot, -1 = fixed
    // set: the direction evaluation, Hessian-action, gather and pullback
    // kernels return at once while *halt != 0 (iterations a device-resident
    // PCG queues past its stop)
    const double* halt = nullptr;
};

// Several blocks under one stacked control vector (reference
// ControlTopoSessionT, src/control_solve.hpp:1585-2837): every block of the
// session is a domain of cells; some carry a net. Block nodes are numbered
// in one concatenated space (block by block, row-major within a block), so
// are the cell samples and the nets' controls (the stacked vector).
struct TopoBlock {
    CtrlGeom g;             // block geometry; the net's tables when net >= 0 (all its controls free)
    long long node0 = 0;    // first node in the concatenated node space
    long long samp0 = 0;    // first sample in the concatenated sample space
    long long ctrl0 = 0;    // first control of the net in the stacked vector
    const double* b = nullptr;  // the net's offset (block node order), null = 0
    int net = -1;
};
struct TopoView {
    const TopoBlock* blk = nullptr;
    int nb = 0;
    const long long* npfx = nullptr;  // [nb + 1] node prefix over blocks
    const long long* spfx = nullptr;  // [nb + 1] sample prefix over blocks
    int n_nets = 0;
    const int* net_blk = nullptr;     // [n_nets] block of each net
    const long long* cpfx = nullptr;  // [n_nets + 1] stacked control prefix
    // every node of a global point shared by several block nodes (block
    // faces, periodic seams) names one owner, the node whose store slot is
    // the point's root; owner[n] = n for owners and unshared nodes
    const int* owner = nullptr;       // [n_all]
    const int* grp_off = nullptr;     // [n_all + 1] the copies an owner owns
    const int* grp_mem = nullptr;
    long long n_all = 0, ns_all = 0, n_stack = 0;
    int dim = 2;
    // set: the evaluation, Hessian-action, gather and pullback kernels
    // return at once while *halt != 0 (the iterations a device-resident
    // PCG queues past its stop)
    const double* halt = nullptr;
};

// FP64 positions of every block node from the FP32 store (its own slot).
void launch_topo_load(const TopoView& t, const float* x, double* out, hipStream_t s);
// Every block node takes its owner's net value: state (src = stacked C;
// nodes whose owner block has no net keep their value) or direction (src =
// a stacked direction; 0 where the owner has no net); the store receives
// the state at each node's own slot when out_plane is given.
void launch_topo_eval(const TopoView& t, bool dir, const double* src, double* out, float* out_plane, hipStream_t s);
// The per-sample and pullback stages of the single-block path over the
// concatenated spaces: samples (gs, hs component-major over ns_all) and the
// per-control pullback into the stacked vector (mode 0: D, mode 1: the
// Jacobi blocks D x D).
void launch_topo_samples(const TopoView& t, int kind, const double* xb, double* gs, double* hs, hipStream_t s);
void launch_topo_pullback(const TopoView& t, int mode, const double* y, double* out, hipStream_t s);
// The owner gather of the D-valued node field in two steps: per sample its
// Hessian action (hv) or metric gradient summed into the cell's corners
// inside the cell's lanes (cc: D values per cell corner, in the sample index
// space), then per node the corners of its incident cells and its copies'
// (copies receive 0).
void launch_topo_cell(const TopoView& t, bool hv, const double* gs, const double* hs, const double* dir, double* cc,
                      hipStream_t s);
void launch_topo_node_cc(const TopoView& t, const double* cc, double* out, hipStream_t s);
// The owners' contracted node Hessians (sum over incident samples of
// u^T H u, D x D; copies 0) through per-cell corner sums (cc: D x D values
// per cell corner).
void launch_topo_node_hessians(const TopoView& t, const double* hs, double* cc, double* out, hipStream_t s);
// Energy and min det over every block's cells at xb + a0 dir and xb + a1 dir
// (as launch_ctrl_energy).
void launch_topo_energy(const TopoView& t, int kind, const double* xb, const double* dir, double a0, double a1,
                        double* partials, double* out, hipStream_t s);
// y = A x + lam z (z null: no term) over n_rows CSR rows.
void launch_csr_apply(long long n_rows, const int* off, const int* col, const double* coef, const double* x,
                      double lam, const double* z, double* y, hipStream_t s);
// y_i += sum_e coef_e w_{row_e} r_{row_e} over transpose row i.
void launch_csrT_accum(long long n_rows, const int* off, const int* row, const double* coef, const double* w,
                       const double* r, double* y, hipStream_t s);
// diag_c = sum over the controls in column c of R (transpose CSR, rows
// ascending) of v^T P_control v; z = r / (diag + lam) (r where not > 0).
void launch_qdiag_build(int dim, long long nq, const int* g_off, const int* g_row, const double* g_coef,
                        const double* P, double* diag, hipStream_t s);
void launch_qdiag_apply(long long nq, const double* diag, double lam, const double* r, double* z, hipStream_t s);

// The PCG recurrences of the control loops with device-resident scalars
// (s, 8 doubles: 0 r.z, 1 r.r, 2 p.Ap, 3 |b|^2, 4 stop threshold, 5
// iterations, 6 state 0 running / 1 converged / 2 negative curvature / 3
// iteration cap), one workgroup each. The preconditioner P: block false, a
// diagonal (nq values, z = r / (P + lam) where positive); block true, one
// dim x dim block per control (nq / dim of them, z = (P_f + lam I)^-1 r_f).
// init: x = 0, r = -G, z, p = z. pap: p.Ap. update (skipped once the state
// is not running): the step, the residual and its norm, the stop test,
// then z, r.z and p = z + beta p.
void launch_qpcg_init(int dim, bool block, long long nq, const double* G, const double* P, double lam, double rtol,
                      double* x, double* r, double* z, double* p, double* s, hipStream_t st);
void launch_qpcg_pap(long long nq, const double* p, const double* Ap, double* s, hipStream_t st);
void launch_qpcg_update(int dim, bool block, long long nq, const double* P, double lam, int max_iter, double* x,
                        double* r, double* z, double* p, const double* Ap, double* s, hipStream_t st);

// Evaluation of the map over the block's nodes into block-local FP64 node
// fields (dim per node, node order row-major) and, when given, into the FP32
// store at the block's node slots:
//   state: M C + b
//   dir:   M v with v a free-control vector (fixed controls 0, no b)
void launch_ctrl_eval_state(const CtrlGeom& g, const double* C, const double* b, double* out, float* out_plane,
                            hipStream_t s);
void launch_ctrl_eval_dir(const CtrlGeom& g, const double* v, double* out, float* out_plane, hipStream_t s);
// Per sample of the block's cells (cell-major, corner bits axis 0 first):
// dmu/dA (dim^2) and the symmetric d2mu/dA2 (upper triangle row-major,
// dim^2 (dim^2 + 1) / 2) in FP64 at the block-local positions xb. kind:
// 0 shape, 2 shape_size.
void launch_ctrl_samples(const CtrlGeom& g, int kind, const double* xb, double* gs, double* hs, hipStream_t s);
// The block's cell energy and min det at xb + a0 dir and xb + a1 dir (dir
// null: at xb), FP64: out = {E(a0), md(a0), E(a1), md(a1)}; partials holds
// 4 kCtrlParts doubles.
constexpr int kCtrlParts = 256;
void launch_ctrl_energy(const CtrlGeom& g, int kind, const double* xb, const double* dir, double a0, double a1,
                        double* partials, double* out, hipStream_t s);
// Per sample: zs (dim^2) = H_p dA_p, dA_p the sample's A of the direction
// field dir (block-local node order).
// Per sample: its Hessian action H_p dA_p (hv; dA_p from the block-local
// direction field dir) or its metric gradient (gs), summed into the cell's
// corners inside the cell's lanes (cc: D values per cell corner, in the
// sample index space); then per block node the corner entries of its
// incident cells: the node field of the energy gradient or of the Hessian
// action, without a per-sample intermediate.
void launch_ctrl_cell(const CtrlGeom& g, bool hv, const double* gs, const double* hs, const double* dir, double* cc,
                      hipStream_t s);
void launch_ctrl_node_cc(const CtrlGeom& g, const double* cc, double* out, hipStream_t s);
// The contracted node Hessians (sum over incident samples of u^T H u with
// u the node's column weights, D x D per node) through per-cell corner sums
// (cc: D x D values per cell corner).
void launch_ctrl_node_hessians(const CtrlGeom& g, const double* hs, double* cc, double* out, hipStream_t s);
// Per free control: mode 0 out (dim) = sum_nodes w * y; mode 1 out (dim^2) =
// sum_nodes w^2 * y (dim^2 blocks: the Jacobi preconditioner).
void launch_ctrl_pullback(const CtrlGeom& g, int mode, const double* y, double* out, hipStream_t s);
// z = (P_f + lam I)^-1 r per free control (dim x dim blocks).
void launch_ctrl_precond(const CtrlGeom& g, const double* P, double lam, const double* r, double* z, hipStream_t s);
// One-workgroup FP64 reductions into *out: 0 x.y, 1 max |x|, 2 the trace sum
// of n dim x dim blocks of x.
void launch_ctrl_reduce(int op, const double* x, const double* y, long long n, int dim, double* out, hipStream_t s);
// y = a x + y (xpay = false) or y = x + a y (xpay = true); n doubles.
void launch_ctrl_axpy(long long n, double a, const double* x, double* y, bool xpay, hipStream_t s);
// y = -x (n doubles)
void launch_ctrl_neg(long long n, const double* x, double* y, hipStream_t s);
// C[free_idx[f]] += a * v[f] (dim values each)
void launch_ctrl_scatter_free(const CtrlGeom& g, double a, const double* v, double* C, hipStream_t s);

}  // namespace eggmi
