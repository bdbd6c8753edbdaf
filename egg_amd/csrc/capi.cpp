// C ABI of the MI355X structured smoother (include/eggmi.h). Converts
// exceptions to status codes and keeps the calling thread's last message.
#include "eggmi.h"
#include "entity.hpp"
#include "mlp_host.hpp"
#include "pcg.hpp"
#include "session.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <cstring>
#include <vector>
#include <new>
#include <string>

struct eggmi_session {
    eggmi::Session s;
};

struct eggmi_mlp {
    std::unique_ptr<eggmi::Mlp> m;
};

namespace
{
thread_local std::string g_last_error;

template <class F> eggmi_status guarded(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return EGGMI_OK;
    } catch (const eggmi::InvalidArgument& e) {
        g_last_error = e.what();
        return EGGMI_INVALID_ARGUMENT;
    } catch (const eggmi::NoDevice& e) {
        g_last_error = e.what();
        return EGGMI_NO_DEVICE;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return EGGMI_INTERNAL;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return EGGMI_INTERNAL;
    }
}

void need(const void* p, const char* what)
{
    if (!p) throw eggmi::InvalidArgument(std::string(what) + " must not be null");
}

int phase_kind(int phase)
{
    if (phase < 0 || phase > 2) throw eggmi::InvalidArgument("unknown phase " + std::to_string(phase));
    return phase;  // eggmi_phase values equal the kernel Kind values
}
}  // namespace

extern "C" {

int eggmi_ping(void) { return 0; }

const char* eggmi_last_error(void) { return g_last_error.c_str(); }

eggmi_status eggmi_session_create(int dim, const eggmi_group* groups, int n_groups, const eggmi_stencil* stencil,
                                  const eggmi_structured* structured, const double* X, int64_t num_nodes,
                                  int device, eggmi_session** out)
{
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        need(stencil, "stencil");
        need(structured, "structured");
        need(X, "X");
        if (n_groups < 0) throw eggmi::InvalidArgument("n_groups must be >= 0");
        if (n_groups > 0) need(groups, "groups");
        auto* s = new eggmi_session;
        try {
            s->s.create(dim, groups, n_groups, stencil, structured, X, num_nodes, device);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

eggmi_status eggmi_session_destroy(eggmi_session* s)
{
    return guarded([&] { delete s; });
}

eggmi_status eggmi_run(eggmi_session* s, int n_sweeps, int phase, double delta, double omega, int report_every,
                       double* energies, double* mindets, int64_t capacity, int64_t* n_reports)
{
    return guarded([&] {
        need(s, "session");
        need(n_reports, "n_reports");
        const int kind = phase_kind(phase);
        const int k = report_every <= 0 ? n_sweeps : report_every;
        const int64_t want = n_sweeps <= 0 ? 0 : (n_sweeps + k - 1) / k;
        if (want > capacity) throw eggmi::InvalidArgument("report capacity too small");
        auto [e, m] = s->s.run_jacobi(n_sweeps, kind, delta, omega, report_every);
        for (size_t i = 0; i < e.size(); ++i) {
            energies[i] = e[i];
            mindets[i] = m[i];
        }
        *n_reports = (int64_t)e.size();
    });
}

eggmi_status eggmi_run_fas(eggmi_session* s, int n_cycles, int nu_pre, int nu_post, int nu_coarse, double omega,
                           int max_levels, int phase, double* energies, double* mindets, int64_t capacity,
                           int64_t* n_reports)
{
    return guarded([&] {
        need(s, "session");
        need(n_reports, "n_reports");
        const int kind = phase_kind(phase);
        if (kind == 1)
            throw eggmi::InvalidArgument("run_fas: FAS smooths the barrier phases only; use run(phase=\"untangle\") "
                                         "to untangle first");
        auto [e, m] = s->s.run_fas(n_cycles, nu_pre, nu_post, nu_coarse, omega, max_levels, kind);
        if ((int64_t)e.size() > capacity) throw eggmi::InvalidArgument("report capacity too small");
        for (size_t i = 0; i < e.size(); ++i) {
            energies[i] = e[i];
            mindets[i] = m[i];
        }
        *n_reports = (int64_t)e.size();
    });
}

eggmi_status eggmi_mg_levels(eggmi_session* s, int max_levels, int32_t* shapes_out, int64_t capacity,
                             int32_t* n_levels, int32_t* n_blocks)
{
    return guarded([&] {
        need(s, "session");
        auto lv = s->s.mg_level_shapes(max_levels);
        const int d = s->s.dim;
        const int64_t need_n = (int64_t)lv.size() * (int64_t)s->s.fine.lay.shape.size() * d;
        if (n_levels) *n_levels = (int32_t)lv.size();
        if (n_blocks) *n_blocks = (int32_t)s->s.fine.lay.shape.size();
        if (!shapes_out) return;
        if (need_n > capacity) throw eggmi::InvalidArgument("shape capacity too small");
        int64_t q = 0;
        for (auto& level : lv)
            for (auto& sh : level)
                for (int k = 0; k < d; ++k) shapes_out[q++] = sh[k];
    });
}

eggmi_status eggmi_mg_level_tables(eggmi_session* s, int level, int64_t* n_slots, int64_t* n_halo,
                                   int64_t* n_share, int32_t* copy_src, int32_t* copy_dst, uint8_t* free_mask,
                                   uint8_t* free_interior)
{
    return guarded([&] {
        need(s, "session");
        s->s.build_hierarchy();
        if (level < 0 || level > (int)s->s.coarse.size())
            throw eggmi::InvalidArgument("mg level " + std::to_string(level) + " outside [0, " +
                                         std::to_string(s->s.coarse.size()) + "]");
        const eggmi::Level& L = level == 0 ? s->s.fine : *s->s.coarse[level - 1];
        if (n_slots) *n_slots = L.lay.n_slots();
        if (n_halo) *n_halo = L.n_halo + L.n_fan;
        if (n_share) *n_share = L.n_share;
        if (copy_src) std::copy(L.copy_src.begin(), L.copy_src.end(), copy_src);
        if (copy_dst) std::copy(L.copy_dst.begin(), L.copy_dst.end(), copy_dst);
        const std::vector<uint8_t>& fr = level == 0 ? s->s.fine_free : L.free_mask;
        const std::vector<uint8_t>& fi = level == 0 ? s->s.fine.mask : L.free_interior;
        if (free_mask) std::copy(fr.begin(), fr.end(), free_mask);
        if (free_interior) std::copy(fi.begin(), fi.end(), free_interior);
    });
}

eggmi_status eggmi_mg_probe(eggmi_session* s, int level, int op, const double* in, double* out)
{
    return guarded([&] {
        need(s, "session");
        need(out, "out");
        if (op != 3) need(in, "in");
        s->s.mg_probe(level, op, in, out);
    });
}

eggmi_status eggmi_geometry_eval(int dim, const double* record, int64_t len, int64_t n, const double* q,
                                 double* proj, double* basis)
{
    return guarded([&] {
        if (dim != 2 && dim != 3) throw eggmi::InvalidArgument("geometry: dim must be 2 or 3");
        if (n < 0) throw eggmi::InvalidArgument("geometry: n must be >= 0");
        need(record, "record");
        if (len < 1) throw eggmi::InvalidArgument("geometry: empty record");
        const int tag = record[0] == 0.0 ? 0 : (int)record[0];
        if (tag != 0) eggmi::entity_blob_check(record, len, dim);
        const int K = tag == 0 ? dim : eggmi::ent_tdim(tag);
        if (n == 0) return;
        need(q, "q");
        for (int64_t i = 0; i < n; ++i) {
            double f[3] = {}, B[3][3] = {};
            if (tag == 0) {  // free: identity, the unit basis
                for (int r = 0; r < dim; ++r) {
                    f[r] = q[dim * i + r];
                    B[r][r] = 1.0;
                }
            } else {
                eggmi::ent_frame(tag, record, q + dim * i, f, B);
            }
            if (proj)
                for (int r = 0; r < dim; ++r) proj[dim * i + r] = f[r];
            if (basis)
                for (int r = 0; r < dim; ++r)
                    for (int c = 0; c < K; ++c) basis[(dim * i + r) * K + c] = B[c][r];
        }
    });
}

eggmi_status eggmi_get_X(eggmi_session* s, double* X_out, int64_t n)
{
    return guarded([&] {
        need(s, "session");
        need(X_out, "X_out");
        if (n < s->s.M * s->s.dim) throw eggmi::InvalidArgument("X_out too small");
        s->s.get_X(X_out);
    });
}

eggmi_status eggmi_set_X(eggmi_session* s, const double* X, int64_t n)
{
    return guarded([&] {
        need(s, "session");
        need(X, "X");
        if (n != s->s.M * s->s.dim) throw eggmi::InvalidArgument("X length must be d*N");
        s->s.set_X(X);
    });
}

eggmi_status eggmi_session_info(eggmi_session* s, eggmi_info* info)
{
    return guarded([&] {
        need(s, "session");
        need(info, "info");
        const auto& L = s->s.fine;
        info->n_slots = L.lay.n_slots();
        info->n_overflow = L.lay.overflow;
        info->n_halo = L.n_halo;
        info->n_fan = L.n_fan;
        info->n_share = L.n_share;
        info->n_tiles = (int64_t)L.tiles.size();
        info->n_stored = (int64_t)L.st_dof.size();
        info->n_synth = L.n_synth;
        info->structured_energy = L.structured_energy ? 1 : 0;
    });
}

eggmi_status eggmi_session_tables(eggmi_session* s, int32_t* copy_src, int32_t* copy_dst, int32_t* g2s,
                                  int64_t* block_base)
{
    return guarded([&] {
        need(s, "session");
        const auto& L = s->s.fine;
        if (copy_src) std::memcpy(copy_src, L.copy_src.data(), L.copy_src.size() * sizeof(int32_t));
        if (copy_dst) std::memcpy(copy_dst, L.copy_dst.data(), L.copy_dst.size() * sizeof(int32_t));
        if (g2s) std::memcpy(g2s, s->s.g2s.data(), s->s.g2s.size() * sizeof(int32_t));
        if (block_base)
            for (size_t b = 0; b < L.lay.base.size(); ++b) block_base[b] = L.lay.base[b];
    });
}

eggmi_status eggmi_get_packed(eggmi_session* s, float* out, int64_t n)
{
    return guarded([&] {
        need(s, "session");
        need(out, "out");
        auto v = s->s.packed();
        if ((int64_t)v.size() > n) throw eggmi::InvalidArgument("out too small");
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    });
}

eggmi_status eggmi_halo_exchange(eggmi_session* s)
{
    return guarded([&] {
        need(s, "session");
        s->s.need_device();
        auto& L = s->s.fine;
        s->s.halo(L, L.x[L.cur].p);
        EGGMI_HIP(hipStreamSynchronize(s->s.stream));
    });
}

eggmi_status eggmi_set_tile_variant(eggmi_session* s, int variant, int* n_variants)
{
    return guarded([&] {
        need(s, "session");
        const int nv = eggmi::num_tile_variants(s->s.dim);
        if (n_variants) *n_variants = nv;
        if (variant < 0) return;
        s->s.set_tile_variant(variant);
    });
}

eggmi_status eggmi_energy(eggmi_session* s, int phase, double delta, double* energy, double* mindet)
{
    return guarded([&] {
        need(s, "session");
        const auto r = s->s.energy(phase_kind(phase), delta);
        if (energy) *energy = r.first;
        if (mindet) *mindet = r.second;
    });
}

eggmi_status eggmi_autotune(eggmi_session* s, int n_sweeps, double omega, double* ms, int capacity, int* best)
{
    return guarded([&] {
        need(s, "session");
        const auto t = s->s.autotune(n_sweeps, omega);
        if (ms)
            for (int v = 0; v < capacity && v < (int)t.size(); ++v) ms[v] = t[v];
        if (best) *best = s->s.tile_variant;
    });
}

eggmi_status eggmi_set_level_variants(eggmi_session* s, const int* variants, int n)
{
    return guarded([&] {
        need(s, "session");
        if (n < 0) throw eggmi::InvalidArgument("n must be >= 0");
        if (n > 0) need(variants, "variants");
        s->s.set_level_variants(std::vector<int>(variants, variants + n));
    });
}

eggmi_status eggmi_set_multi_sweep(eggmi_session* s, int level, int k, int ty, int tx, int kernel, int* n_kernels)
{
    return guarded([&] {
        need(s, "session");
        if (n_kernels) *n_kernels = eggmi::num_ms_variants();
        s->s.set_multi_sweep(level, k, ty, tx, kernel);
    });
}

eggmi_status eggmi_multi_sweep_info(eggmi_session* s, int level, int64_t* info, double* check_err, unsigned seed)
{
    return guarded([&] {
        need(s, "session");
        const eggmi::Level& L = s->s.level_ref(level);
        if (info) {
            const int64_t v[8] = {L.ms_sweeps,
                                  L.ms.ty,
                                  L.ms.tx,
                                  L.ms_kernel,
                                  (int64_t)L.ms.tiles.size(),
                                  L.ms.lds_stride,
                                  L.ms.core_dofs,
                                  L.ms.updates};
            std::copy(v, v + 8, info);
        }
        if (check_err) {  // every launch depth the tables serve (cadence remainders run fewer sweeps)
            double e = L.multi_selfcheck(L.ms.k, seed);
            for (int k = 1; k < L.ms.k; ++k) e = std::max(e, L.multi_selfcheck(k, seed + (unsigned)k));
            *check_err = e;
        }
    });
}

eggmi_status eggmi_autotune_multi(eggmi_session* s, int n_sweeps, double omega, const int32_t* cand, int n_cands,
                                  double base_ms, double* ms)
{
    return guarded([&] {
        need(s, "session");
        if (n_cands < 0) throw eggmi::InvalidArgument("n_cands must be >= 0");
        if (n_cands > 0) need(cand, "cand");
        std::vector<std::array<int, 4>> c(n_cands);
        for (int i = 0; i < n_cands; ++i) c[i] = {cand[4 * i], cand[4 * i + 1], cand[4 * i + 2], cand[4 * i + 3]};
        const auto t = s->s.autotune_multi(n_sweeps, omega, c, base_ms);
        if (ms) std::copy(t.begin(), t.end(), ms);
    });
}

eggmi_status eggmi_autotune_levels_multi(eggmi_session* s, int n_sweeps, double omega, int max_levels, int phase,
                                         const int32_t* cand, int n_cands, double* ms, int64_t capacity,
                                         int* n_levels)
{
    return guarded([&] {
        need(s, "session");
        need(n_levels, "n_levels");
        if (n_cands < 0) throw eggmi::InvalidArgument("n_cands must be >= 0");
        if (n_cands > 0) need(cand, "cand");
        std::vector<std::array<int, 4>> c(n_cands);
        for (int i = 0; i < n_cands; ++i) c[i] = {cand[4 * i], cand[4 * i + 1], cand[4 * i + 2], cand[4 * i + 3]};
        const auto t = s->s.autotune_levels_multi(n_sweeps, omega, max_levels, phase_kind(phase), c);
        if (ms)
            for (int64_t i = 0; i < capacity && i < (int64_t)t.size(); ++i) ms[i] = t[i];
        *n_levels = (int)(t.size() / (size_t)(n_cands + 1));
    });
}

eggmi_status eggmi_autotune_levels(eggmi_session* s, int n_sweeps, double omega, int max_levels, int phase,
                                   double* ms, int64_t capacity, int* best, int best_capacity, int* n_levels)
{
    return guarded([&] {
        need(s, "session");
        need(n_levels, "n_levels");
        const auto t = s->s.autotune_levels(n_sweeps, omega, max_levels, phase_kind(phase));
        const int nv = 2 * eggmi::num_tile_variants(s->s.dim);
        const int nl = (int)(t.size() / nv);
        if (ms)
            for (int64_t i = 0; i < capacity && i < (int64_t)t.size(); ++i) ms[i] = t[i];
        if (best)
            for (int i = 0; i < best_capacity && i < nl; ++i) best[i] = s->s.coarse_variant((size_t)i);
        *n_levels = nl;
    });
}

eggmi_status eggmi_last_timing(eggmi_session* s, double* ms, int64_t* sweep_launches, double* sweep_kernel_ms)
{
    return guarded([&] {
        need(s, "session");
        if (ms) *ms = s->s.last_ms;
        if (sweep_launches) *sweep_launches = s->s.last_launches;
        if (sweep_kernel_ms) *sweep_kernel_ms = s->s.last_kernel_ms;
    });
}

eggmi_status eggmi_fas_cycle_kernels(eggmi_session* s, int64_t* kernels)
{
    return guarded([&] {
        need(s, "session");
        need(kernels, "kernels");
        *kernels = s->s.last_fas_graph_kernels;
    });
}

eggmi_status eggmi_set_kernel_timing(eggmi_session* s, int on)
{
    return guarded([&] {
        need(s, "session");
        s->s.kernel_timing = on != 0;
    });
}

eggmi_status eggmi_set_graph_chunk(eggmi_session* s, int chunk)
{
    return guarded([&] {
        need(s, "session");
        s->s.need_device();
        if (chunk < 0 || chunk % 2 != 0) throw eggmi::InvalidArgument("graph chunk must be an even number >= 0");
        s->s.release_graphs();
        s->s.graph_chunk = chunk;
    });
}

eggmi_status eggmi_set_counters(eggmi_session* s, int on)
{
    return guarded([&] {
        need(s, "session");
        s->s.counting = on != 0;
    });
}

eggmi_status eggmi_last_counters(eggmi_session* s, int64_t* out, int n_levels)
{
    return guarded([&] {
        need(s, "session");
        need(out, "out");
        if (n_levels < 0) throw eggmi::InvalidArgument("n_levels must be >= 0");
        const auto& c = s->s.last_counters;
        constexpr int K = eggmi::Session::kCountersPerLevel;
        for (int l = 0; l < n_levels; ++l)
            for (int j = 0; j < 3; ++j) {
                const size_t i = (size_t)l * K + j;
                out[3 * l + j] = i < c.size() ? (int64_t)c[i] : 0;
            }
    });
}

eggmi_status eggmi_mlp_create(int dim, uint64_t seed, int64_t capacity, eggmi_mlp** out)
{
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        auto h = std::make_unique<eggmi_mlp>();
        h->m = std::make_unique<eggmi::Mlp>(dim, seed, capacity);
        *out = h.release();
    });
}

eggmi_status eggmi_mlp_destroy(eggmi_mlp* m)
{
    return guarded([&] { delete m; });
}

eggmi_status eggmi_mlp_shape(eggmi_mlp* m, int32_t* shape, int64_t* n_params, int64_t* n_data, int64_t* capacity)
{
    return guarded([&] {
        need(m, "mlp");
        const auto& n = *m->m;
        if (shape) {
            shape[0] = n.dim;
            shape[1] = n.in;
            shape[2] = n.inp;
            shape[3] = n.hidden;
            shape[4] = n.out;
        }
        if (n_params) *n_params = n.n_params;
        if (n_data) *n_data = n.n;
        if (capacity) *capacity = n.cap;
    });
}

eggmi_status eggmi_mlp_get_params(eggmi_mlp* m, float* out, int64_t n)
{
    return guarded([&] {
        need(m, "mlp");
        need(out, "out");
        if (n != m->m->n_params) throw eggmi::InvalidArgument("params length must be " + std::to_string(m->m->n_params));
        m->m->get_params(out);
    });
}

eggmi_status eggmi_mlp_set_params(eggmi_mlp* m, const float* in, int64_t n)
{
    return guarded([&] {
        need(m, "mlp");
        need(in, "params");
        if (n != m->m->n_params) throw eggmi::InvalidArgument("params length must be " + std::to_string(m->m->n_params));
        m->m->set_params(in);
    });
}

eggmi_status eggmi_mlp_forward(eggmi_mlp* m, const float* X, int64_t rows, float* Y)
{
    return guarded([&] {
        need(m, "mlp");
        if (rows > 0) {
            need(X, "X");
            need(Y, "Y");
        }
        m->m->forward(X, rows, Y);
    });
}

eggmi_status eggmi_mlp_data_add(eggmi_mlp* m, const float* X, const float* T, int64_t rows)
{
    return guarded([&] {
        need(m, "mlp");
        if (rows > 0) {
            need(X, "X");
            need(T, "T");
        }
        m->m->data_add(X, T, rows);
    });
}

eggmi_status eggmi_mlp_data_get(eggmi_mlp* m, float* X, float* T, int64_t offset, int64_t rows)
{
    return guarded([&] {
        need(m, "mlp");
        m->m->data_get(X, T, offset, rows);
    });
}

eggmi_status eggmi_mlp_data_clear(eggmi_mlp* m)
{
    return guarded([&] {
        need(m, "mlp");
        m->m->data_clear();
    });
}

eggmi_status eggmi_mlp_gradient(eggmi_mlp* m, const int32_t* rows_idx, int64_t batch, float* grad, double* loss)
{
    return guarded([&] {
        need(m, "mlp");
        need(rows_idx, "rows_idx");
        need(grad, "grad");
        const double l = m->m->gradient(rows_idx, batch, grad);
        if (loss) *loss = l;
    });
}

eggmi_status eggmi_mlp_train(eggmi_mlp* m, int epochs, int64_t batch, float lr, uint64_t seed, double* epoch_loss)
{
    return guarded([&] {
        need(m, "mlp");
        if (epochs < 0) throw eggmi::InvalidArgument("epochs must be >= 0");
        const auto h = m->m->train(epochs, batch, lr, seed);
        if (epoch_loss) std::copy(h.begin(), h.end(), epoch_loss);
    });
}

eggmi_status eggmi_session_harvest(eggmi_session* s, eggmi_mlp* m, int n_snap, int every, int horizon,
                                   double omega, int64_t* added)
{
    return guarded([&] {
        need(s, "session");
        need(m, "mlp");
        const long long a = eggmi::harvest(s->s, *m->m, n_snap, every, horizon, omega);
        if (added) *added = a;
    });
}

eggmi_status eggmi_session_set_proposal(eggmi_session* s, eggmi_mlp* m, int mode, float scale)
{
    return guarded([&] {
        need(s, "session");
        s->s.need_device();
        if (m && m->m->dim != s->s.dim) throw eggmi::InvalidArgument("proposal network dimension differs from the session's");
        if (mode != 0 && mode != 1) throw eggmi::InvalidArgument("proposal mode must be 0 (replace) or 1 (predict-correct)");
        if (!(scale > 0.0f) || !std::isfinite(scale)) throw eggmi::InvalidArgument("proposal scale must be positive");
        if (m && eggmi::tile_variant_proposal_partner(s->s.dim, s->s.tile_variant) < 0)
            s->s.set_tile_variant(0);  // no proposal kernel runs on this variant's tile table
        s->s.proposal = m ? m->m.get() : nullptr;
        s->s.proposal_mode = mode;
        s->s.proposal_scale = scale;
    });
}

eggmi_status eggmi_session_set_control(eggmi_session* s, const eggmi_control* c)
{
    return guarded([&] {
        need(s, "session");
        need(c, "control");
        s->s.set_control(*c);
    });
}

eggmi_status eggmi_session_set_control_topo(eggmi_session* s, const eggmi_control_topo* t)
{
    return guarded([&] {
        need(s, "session");
        need(t, "control topology");
        s->s.set_control_topo(*t);
    });
}

eggmi_status eggmi_control_set_reduction(eggmi_session* s, int64_t nq, const int32_t* r_off, const int32_t* r_col,
                                         const double* r_coef)
{
    return guarded([&] {
        need(s, "session");
        s->s.control_topo_set_reduction(nq, r_off, r_col, r_coef);
    });
}

eggmi_status eggmi_control_set_penalty(eggmi_session* s, int64_t n_pen, const int32_t* p_off, const int32_t* p_col,
                                       const double* p_coef, const double* p_w)
{
    return guarded([&] {
        need(s, "session");
        s->s.control_topo_set_penalty(n_pen, p_off, p_col, p_coef, p_w);
    });
}

eggmi_status eggmi_control_set_b(eggmi_session* s, int32_t net, const double* b, int64_t n)
{
    return guarded([&] {
        need(s, "session");
        s->s.control_topo_set_b(net, b, n);
    });
}

eggmi_status eggmi_control_set_C(eggmi_session* s, const double* C, int64_t n)
{
    return guarded([&] {
        need(s, "session");
        need(C, "C");
        s->s.control_set_C(C, n);
    });
}

eggmi_status eggmi_control_get_C(eggmi_session* s, double* C, int64_t n)
{
    return guarded([&] {
        need(s, "session");
        need(C, "C");
        s->s.control_get_C(C, n);
    });
}

eggmi_status eggmi_run_control(eggmi_session* s, const eggmi_control_options* o, int phase, double* energies,
                               double* mindets, double* alphas, int64_t capacity, int64_t* n_reports,
                               double* pcg_iters, double* pcg_relres, int64_t pcg_capacity, int64_t* n_pcg,
                               int32_t* iters, int32_t* converged)
{
    return guarded([&] {
        need(s, "session");
        need(o, "options");
        const int kind = phase_kind(phase);
        if (o->n_outer < 0 || o->lm_tries < 1 || o->pcg_max_iter < 0)
            throw eggmi::InvalidArgument("run_control: n_outer, pcg_max_iter must be >= 0 and lm_tries >= 1");
        if (!(o->alpha_min > 0.0) || !(o->pcg_rtol >= 0.0) || !(o->grad_tol >= 0.0) || !(o->lm0 >= 0.0))
            throw eggmi::InvalidArgument("run_control: alpha_min must be > 0; grad_tol, pcg_rtol, lm0 >= 0");
        eggmi::ControlOptions op;
        op.n_outer = o->n_outer;
        op.grad_tol = o->grad_tol;
        op.alpha_min = o->alpha_min;
        op.lm0 = o->lm0;
        op.lm_tries = o->lm_tries;
        op.pcg_max_iter = o->pcg_max_iter;
        op.pcg_rtol = o->pcg_rtol;
        op.pcg_forcing = o->pcg_forcing != 0;
        const eggmi::ControlReport r = s->s.run_control(op, kind);
        const int64_t nr = (int64_t)r.energies.size(), np = (int64_t)r.pcg_iters.size();
        for (int64_t i = 0; i < nr && i < capacity; ++i) {
            if (energies) energies[i] = r.energies[i];
            if (mindets) mindets[i] = r.mindets[i];
            if (alphas && i + 1 < nr) alphas[i] = r.alphas[i];
        }
        for (int64_t i = 0; i < np && i < pcg_capacity; ++i) {
            if (pcg_iters) pcg_iters[i] = r.pcg_iters[i];
            if (pcg_relres) pcg_relres[i] = r.pcg_relres[i];
        }
        if (n_reports) *n_reports = nr;
        if (n_pcg) *n_pcg = np;
        if (iters) *iters = r.iters;
        if (converged) *converged = r.converged ? 1 : 0;
    });
}

eggmi_status eggmi_session_set_proposal_every(eggmi_session* s, int every)
{
    return guarded([&] {
        need(s, "session");
        if (every < 1) throw eggmi::InvalidArgument("proposal cadence must be >= 1");
        s->s.proposal_every = every;
    });
}

}  // extern "C"

namespace
{
// One solve on a stream of its own: the work area in one allocation, the
// device time of the solve (uploads and downloads excluded) in *solve_ms.
template <class Solve>
void pcg_on_device(int64_t n, const double* b, double* x, double* solve_ms, Solve&& run)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw eggmi::NoDevice("pcg: no HIP device");
    hipStream_t s;
    EGGMI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        eggmi::DArr<double> d_b, d_x, d_work;
        d_b.upload(std::vector<double>(b, b + n));
        d_x.alloc(std::max<int64_t>(n, 1));
        d_work.alloc(eggmi::pcg_work_doubles(n));
        eggmi::PcgWork w;
        eggmi::pcg_carve(d_work.p, n, w);
        EGGMI_HIP(hipEventCreate(&e0));
        EGGMI_HIP(hipEventCreate(&e1));
        EGGMI_HIP(hipEventRecord(e0, s));
        run(d_b.p, d_x.p, w, s);
        EGGMI_HIP(hipEventRecord(e1, s));
        EGGMI_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        EGGMI_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (solve_ms) *solve_ms = ms;
        if (n > 0) EGGMI_HIP(hipMemcpy(x, d_x.p, n * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    EGGMI_HIP(hipStreamDestroy(s));
}

eggmi::PcgRecurrence pcg_recurrence(int recurrence)
{
    if (recurrence != 0 && recurrence != 1)
        throw eggmi::InvalidArgument("pcg: recurrence must be 0 (standard) or 1 (one reduction per iteration)");
    return recurrence == 0 ? eggmi::PcgRecurrence::kStandard : eggmi::PcgRecurrence::kOneReduction;
}

void pcg_common_checks(int64_t n, const double* b, const double* x, double rtol, int max_iter)
{
    if (n < 0) throw eggmi::InvalidArgument("pcg: n must be >= 0");
    need(b, "b");
    need(x, "x");
    if (max_iter < 0) throw eggmi::InvalidArgument("pcg: max_iter must be >= 0");
    if (!(rtol >= 0.0)) throw eggmi::InvalidArgument("pcg: rtol must be >= 0");
}

void pcg_report(const eggmi::PcgResult& res, int* iters, double* relres, int* status)
{
    if (iters) *iters = res.iters;
    if (relres) *relres = res.relres;
    if (status) *status = res.status;
}


}  // namespace

extern "C" {

eggmi_status eggmi_pcg_csr(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                           const double* b, double* x, double rtol, int max_iter, int recurrence, int* iters,
                           double* relres, int* status, double* solve_ms)
{
    return guarded([&] {
        pcg_common_checks(n, b, x, rtol, max_iter);
        const eggmi::PcgRecurrence rec = pcg_recurrence(recurrence);
        need(rowptr, "rowptr");
        const int64_t nnz = n ? rowptr[n] : 0;
        if (rowptr[0] != 0 || nnz < 0) throw eggmi::InvalidArgument("pcg: rowptr must start at 0");
        for (int64_t i = 0; i < n; ++i)
            if (rowptr[i + 1] < rowptr[i]) throw eggmi::InvalidArgument("pcg: rowptr must be non-decreasing");
        bool has_diag_all = true;
        for (int64_t i = 0; i < n && has_diag_all; ++i) {
            bool has = false;
            for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
                if (col[q] < 0 || col[q] >= n) throw eggmi::InvalidArgument("pcg: column index out of range");
                if (col[q] == i && val[q] != 0.0) has = true;
            }
            has_diag_all = has;
        }
        if (!has_diag_all) throw eggmi::InvalidArgument("pcg: every row needs a nonzero diagonal (Jacobi)");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw eggmi::NoDevice("pcg: no HIP device");
        eggmi::DArr<long long> d_rp;
        eggmi::DArr<int> d_col;
        eggmi::DArr<double> d_val;
        d_rp.upload(std::vector<long long>(rowptr, rowptr + n + 1));
        d_col.upload(std::vector<int>(col, col + nnz));
        d_val.upload(std::vector<double>(val, val + nnz));
        const eggmi::CsrView A {n, d_rp.p, d_col.p, d_val.p};
        eggmi::PcgResult res;
        pcg_on_device(n, b, x, solve_ms, [&](const double* db, double* dx, eggmi::PcgWork& w, hipStream_t s) {
            res = eggmi::pcg_solve_csr(A, db, dx, rtol, max_iter, w, s, eggmi::PcgLaunch::kQueued, rec);
        });
        pcg_report(res, iters, relres, status);
    });
}

eggmi_status eggmi_pcg_lattice(int dim, const int32_t* shape, double diag, double off, const double* b, double* x,
                               double rtol, int max_iter, int recurrence, int* iters, double* relres, int* status,
                               double* solve_ms)
{
    return guarded([&] {
        const eggmi::PcgRecurrence rec = pcg_recurrence(recurrence);
        if (dim < 1 || dim > 3) throw eggmi::InvalidArgument("pcg: lattice dim must be 1, 2 or 3");
        need(shape, "shape");
        eggmi::LatticeView A;
        A.dim = dim;
        A.n = 1;
        for (int k = 0; k < dim; ++k) {
            if (shape[k] < 1) throw eggmi::InvalidArgument("pcg: lattice shape entries must be >= 1");
            A.shape[k] = shape[k];
            A.n *= shape[k];
        }
        if (A.n >= (1ll << 31)) throw eggmi::InvalidArgument("pcg: lattice of 2^31 or more unknowns");
        if (!(diag != 0.0) || !std::isfinite(diag) || !std::isfinite(off))
            throw eggmi::InvalidArgument("pcg: lattice diag must be finite and nonzero (Jacobi), off finite");
        A.diag = diag;
        A.off = off;
        pcg_common_checks(A.n, b, x, rtol, max_iter);
        eggmi::PcgResult res;
        pcg_on_device(A.n, b, x, solve_ms, [&](const double* db, double* dx, eggmi::PcgWork& w, hipStream_t s) {
            res = eggmi::pcg_solve_lattice(A, db, dx, rtol, max_iter, w, s, eggmi::PcgLaunch::kQueued, rec);
        });
        pcg_report(res, iters, relres, status);
    });
}

eggmi_status eggmi_pcg_csr_noalloc(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                                   const double* b, double* x, double rtol, int max_iter, int recurrence,
                                   void* stream, int* iters, double* relres, int* status)
{
    return guarded([&] {
        if (n < 0 || n > eggmi::kPcgNoAllocMaxN)
            throw eggmi::InvalidArgument("pcg no-alloc: n = " + std::to_string(n) + " outside [0, " +
                                         std::to_string(eggmi::kPcgNoAllocMaxN) + "], the work area's size");
        if (max_iter < 0) throw eggmi::InvalidArgument("pcg: max_iter must be >= 0");
        if (!(rtol >= 0.0)) throw eggmi::InvalidArgument("pcg: rtol must be >= 0");
        const eggmi::PcgRecurrence rec = pcg_recurrence(recurrence);
        if (n > 0) {
            need(rowptr, "rowptr");
            need(col, "col");
            need(val, "val");
            need(b, "b");
            need(x, "x");
        }
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw eggmi::NoDevice("pcg: no HIP device");
        const eggmi::CsrView A {n, reinterpret_cast<const long long*>(rowptr), col, val};
        const eggmi::PcgResult res =
          eggmi::pcg_solve_csr_noalloc(A, b, x, rtol, max_iter, static_cast<hipStream_t>(stream), rec);
        pcg_report(res, iters, relres, status);
    });
}

}  // extern "C"
