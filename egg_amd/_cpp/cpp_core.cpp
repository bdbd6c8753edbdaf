// pybind11 module egg_amd._cpp.cpp_core: the reference's egg._cpp.cpp_core
// Python surface (reference src/bindings.cpp:1541-2130, stub
// egg/_cpp/cpp_core.pyi:27-68) over the eggmi C ABI (include/eggmi.h).
// Dict unpacking mirrors unpack_context (src/bindings.cpp:334-490); errors
// map EGGMI_INVALID_ARGUMENT -> ValueError, everything else -> RuntimeError.
#include "eggmi.h"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace py = pybind11;

namespace
{

using IArr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;

void check(eggmi_status st)
{
    if (st == EGGMI_OK) return;
    const std::string msg = eggmi_last_error();
    if (st == EGGMI_INVALID_ARGUMENT) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

int phase_id(const std::string& phase)
{
    if (phase == "untangle") return EGGMI_PHASE_UNTANGLE;
    if (phase == "shape_size") return EGGMI_PHASE_SHAPE_SIZE;
    return EGGMI_PHASE_BARRIER;  // "barrier" and any other name select the shape barrier
}

int device_id(const std::string& device)
{
    if (device == "auto") return 0;
    if (device == "gpu") return 1;
    if (device == "cpu") return 2;
    if (device == "tables") return 3;  // host-side index maps only (inspection)
    throw py::value_error("device must be 'auto', 'cpu', 'gpu' or 'tables'; got '" + device + "'");
}

int recurrence_id(const std::string& r)
{
    if (r == "standard") return 0;
    if (r == "one_reduction") return 1;
    throw py::value_error("recurrence must be 'standard' or 'one_reduction'; got '" + r + "'");
}

int require_dim(int d)
{
    if (d != 2 && d != 3) throw py::value_error("dim must be 2 or 3; got " + std::to_string(d));
    return d;
}

size_t checked_num_nodes(const DArr& X, int d)
{
    if (X.ndim() != 1)
        throw py::value_error("X must be a flat 1-D array of length d*N; got ndim=" + std::to_string(X.ndim()) +
                              " (did you pass an (N, d) array? ravel() it first).");
    if (X.shape(0) % d != 0)
        throw py::value_error("X length must be a multiple of d (" + std::to_string(d) +
                              " coordinates per node); got " + std::to_string(X.shape(0)) + ".");
    return (size_t)X.shape(0) / d;
}

IArr get_i(const py::dict& d, const char* key)
{
    if (!d.contains(key)) throw py::value_error(std::string("missing wire field '") + key + "'");
    return d[key].cast<IArr>();
}
DArr get_d(const py::dict& d, const char* key)
{
    if (!d.contains(key)) throw py::value_error(std::string("missing wire field '") + key + "'");
    return d[key].cast<DArr>();
}

// Per-sample energy weights: one value per sample, or a single value shared
// by every sample (reference src/bindings.cpp:247,364-366 reads that one with
// stride 0); the shared form is expanded here so the C ABI sees one per sample.
DArr sample_weights(DArr wt, int64_t n, const char* where)
{
    if ((int64_t)wt.size() == n) return wt;
    if (wt.size() == 1) {
        DArr full((py::ssize_t)n);
        std::fill(full.mutable_data(), full.mutable_data() + n, wt.data()[0]);
        return full;
    }
    throw py::value_error(std::string("validate_context: ") + where + " weight size " + std::to_string(wt.size()) +
                          " != " + std::to_string(n) + " (one per sample, or a single shared value)");
}

// ---- entities: the reference wire's per-kind records and CSR payloads
// (egg/geometry/entity_soa.py:41-60 field table, :598-688 grouping;
// src/geometry.hpp:1717-2520) as the compute core's self-describing records
// (egg_amd/csrc/entity.hpp: tag, fields, payloads).

// SoA fields per tag on the wire (kFields)
constexpr int kWireFields[17] = {0, 4, 3, 4, 10, 9, 6, 8, 8, 10, 6, 2, 10, 8, 9, 4, 6};
// positional parameters of a composite's segment record (kParamPad)
constexpr int kParamPad = 12;
constexpr int kEntBSplineTag = 10;

struct Slice {
    const double* p = nullptr;
    long long n = 0;
};

Slice sub(const Slice& s, long long off, long long n, const char* what)
{
    if (off < 0 || n < 0 || off + n > s.n) throw py::value_error(std::string("wire: ") + what + " outside its payload");
    return {s.p + off, n};
}

long long wire_count(double v, const char* what)
{
    if (!std::isfinite(v) || v < 0 || v != std::floor(v) || v > 1e9)
        throw py::value_error(std::string("wire: ") + what + " must be a count");
    return (long long)v;
}

// B-spline curve record: p n t0 t1 closed has_w | knots ctrl (w)
void bspline_record(double p, double n, double t0, double t1, double closed, double has_w, const Slice& knots,
                    const Slice& ctrl, const Slice& w, std::vector<double>& out)
{
    const long long pp = wire_count(p, "B-spline degree"), nn = wire_count(n, "B-spline control count");
    if (knots.n != nn + pp + 1 || ctrl.n != 2 * nn || (has_w != 0.0 && w.n != nn))
        throw py::value_error("wire: B-spline payload sizes do not match its degree and control count");
    out.insert(out.end(), {double(kEntBSplineTag), p, n, t0, t1, closed, has_w});
    out.insert(out.end(), knots.p, knots.p + knots.n);
    out.insert(out.end(), ctrl.p, ctrl.p + ctrl.n);
    if (has_w != 0.0) out.insert(out.end(), w.p, w.p + w.n);
}

// one segment of a composite from its positional record [tag, params]
// (egg/geometry/entity_encoding.py:82-178); a B-spline's payload offsets
// point into the composite's arena slice
void segment_record(const double* rec, const Slice& arena, std::vector<double>& out)
{
    const double tv = rec[0];
    const int t = (int)tv;
    const double* q = rec + 1;
    if (tv != t || !(t == 1 || t == 2 || t == 3 || t == 15 || (t >= 6 && t <= 10)))
        throw py::value_error("wire: a composite segment must be a 2D curve (tag " + std::to_string(tv) + ")");
    if (t == kEntBSplineTag) {
        const long long nn = wire_count(q[1], "B-spline control count"), pp = wire_count(q[0], "B-spline degree");
        const bool rat = q[7] != 0.0;
        bspline_record(q[0], q[1], q[4], q[5], 0.0, q[7],
                       sub(arena, wire_count(q[2], "knot offset"), nn + pp + 1, "B-spline knots"),
                       sub(arena, wire_count(q[3], "control offset"), 2 * nn, "B-spline control net"),
                       rat ? sub(arena, wire_count(q[6], "weight offset"), nn, "B-spline weights") : Slice {}, out);
        return;
    }
    out.push_back(tv);
    out.insert(out.end(), q, q + kWireFields[t]);
}

// one DOF's record from its kind's fields r[kWireFields[tag]] and payloads
void wire_entity(int tag, const double* r, const std::vector<Slice>& seg, std::vector<double>& out)
{
    auto slot = [&](size_t i) { return i < seg.size() ? seg[i] : Slice {}; };
    switch (tag) {
    case 4:  // sphere: c(3) r ax(3) ay(3) -> + az = ax x ay
    case 12: {  // cylinder: o(3) ax(3) ay(3) r -> o ax ay az r
        const double* ax = tag == 4 ? r + 4 : r + 3;
        const double* ay = ax + 3;
        const double az[3] = {ax[1] * ay[2] - ax[2] * ay[1], ax[2] * ay[0] - ax[0] * ay[2],
                              ax[0] * ay[1] - ax[1] * ay[0]};
        out.push_back(tag);
        if (tag == 4) {
            out.insert(out.end(), r, r + 10);
            out.insert(out.end(), az, az + 3);
        } else {
            out.insert(out.end(), r, r + 9);
            out.insert(out.end(), az, az + 3);
            out.push_back(r[9]);
        }
        return;
    }
    case 10:
        bspline_record(r[0], r[1], r[2], r[3], r[4], r[5], slot(0), slot(1), slot(2), out);
        return;
    case 11: {  // composite: [n_segs, rec_off] + its arena slice
        const Slice arena = slot(0);
        const long long m = wire_count(r[0], "composite segment count"), ro = wire_count(r[1], "record offset");
        if (m < 1) throw py::value_error("wire: a composite needs at least one segment");
        const Slice recs = sub(arena, ro, m * (1 + kParamPad), "composite segment records");
        const size_t base = out.size();
        out.push_back(11);
        out.push_back((double)m);
        out.resize(out.size() + m, 0.0);
        for (long long s = 0; s < m; ++s) {
            out[base + 2 + s] = (double)(out.size() - base);
            segment_record(recs.p + s * (1 + kParamPad), arena, out);
        }
        return;
    }
    case 14: {  // B-spline surface: pu pv nu nv - - - - has_w + 6 payloads
        const long long pu = wire_count(r[0], "surface degree"), pv = wire_count(r[1], "surface degree");
        const long long nu = wire_count(r[2], "surface control count"), nv = wire_count(r[3], "surface control count");
        const bool rat = r[8] != 0.0;
        const Slice ku = slot(0), kv = slot(1), ctrl = slot(2), w = slot(3), verts = slot(4), loops = slot(5);
        if (ku.n != nu + pu + 1 || kv.n != nv + pv + 1 || ctrl.n != 3 * nu * nv || (rat && w.n != nu * nv))
            throw py::value_error("wire: B-spline surface payload sizes do not match its degrees and net");
        if (loops.n >= 2 && verts.n != 2 * wire_count(loops.p[loops.n - 1], "trim loop offset"))
            throw py::value_error("wire: trim vertices do not match the loop table");
        out.insert(out.end(), {14.0, r[0], r[1], r[2], r[3], r[8], (double)loops.n});
        out.insert(out.end(), ku.p, ku.p + ku.n);
        out.insert(out.end(), kv.p, kv.p + kv.n);
        out.insert(out.end(), ctrl.p, ctrl.p + ctrl.n);
        if (rat) out.insert(out.end(), w.p, w.p + w.n);
        out.insert(out.end(), loops.p, loops.p + loops.n);
        if (loops.n >= 2) out.insert(out.end(), verts.p, verts.p + verts.n);
        return;
    }
    default:
        out.push_back(tag);
        out.insert(out.end(), r, r + kWireFields[tag]);
    }
}

// Keeps the converted arrays alive while the C structs point into them.
struct Wire {
    int dim = 2;
    std::vector<IArr> iarrs;
    std::vector<DArr> darrs;
    std::vector<eggmi_group> groups;
    eggmi_stencil es {};
    eggmi_structured st {};
    std::vector<IArr> bgd_arrs;
    std::vector<const int32_t*> bgd_ptrs;
    std::vector<int32_t> shapes;
    IArr hsb, hsp, hdb, hdp, sb, sl;
    std::vector<std::vector<int64_t>> ent_offs;
    std::vector<std::vector<double>> ent_datas;

    const int32_t* keep(IArr a)
    {
        iarrs.push_back(a);
        return iarrs.back().data();
    }
    const double* keep(DArr a)
    {
        darrs.push_back(a);
        return darrs.back().data();
    }

    Wire(const py::dict& ctx, const py::dict& structured, int d) : dim(d)
    {
        const int dd = d * d;
        py::list gl = ctx["groups"].cast<py::list>();
        for (auto item : gl) {
            py::dict g = item.cast<py::dict>();
            eggmi_group eg {};
            eg.ndof = g["D"].cast<int64_t>();
            IArr dof = get_i(g, "dof_idx"), P = get_i(g, "P_of"), role = get_i(g, "role"), gc = get_i(g, "gc");
            if ((int64_t)dof.size() != eg.ndof)
                throw py::value_error("validate_context: group dof_idx size " + std::to_string(dof.size()) +
                                      " != " + std::to_string(eg.ndof));
            if ((int64_t)P.size() != eg.ndof)
                throw py::value_error("validate_context: group P_of size " + std::to_string(P.size()) +
                                      " != " + std::to_string(eg.ndof));
            eg.dof_idx = keep(dof);
            eg.P_of = keep(P);
            eg.total_samples = (int64_t)gc.size();
            if ((int64_t)role.size() != eg.total_samples)
                throw py::value_error("validate_context: group role size " + std::to_string(role.size()) +
                                      " != " + std::to_string(eg.total_samples));
            eg.role = keep(role);
            eg.gc = keep(gc);
            for (int k = 0; k < d; ++k) {
                const std::string gk = "gn" + std::to_string(k), sk = "s" + std::to_string(k);
                IArr gn = get_i(g, gk.c_str());
                DArr s = get_d(g, sk.c_str());
                if ((int64_t)gn.size() != eg.total_samples || (int64_t)s.size() != eg.total_samples)
                    throw py::value_error("validate_context: group " + gk + "/" + sk + " size != " +
                                          std::to_string(eg.total_samples));
                eg.gn[k] = keep(gn);
                eg.s[k] = keep(s);
            }
            DArr W = get_d(g, "W_inv");
            if (W.size() % dd != 0) throw py::value_error("validate_context: group W_inv size not a multiple of d*d");
            eg.w_rows = (int64_t)W.size() / dd;
            eg.W_inv = keep(W);
            if (g.contains("interior_block")) {
                IArr ib = get_i(g, "interior_block"), il = get_i(g, "interior_logical");
                if ((int64_t)ib.size() != eg.ndof || (int64_t)il.size() != eg.ndof * d)
                    throw py::value_error("validate_context: interior_block/interior_logical size mismatch");
                eg.interior_block = keep(ib);
                eg.interior_logical = keep(il);
            } else {
                std::vector<int32_t> neg(eg.ndof, -1), zero(eg.ndof * d, 0);
                eg.interior_block = keep(IArr((py::ssize_t)neg.size(), neg.data()));
                eg.interior_logical = keep(IArr((py::ssize_t)zero.size(), zero.data()));
            }
            if (g.contains("weight") && !g["weight"].is_none())
                eg.weight = keep(sample_weights(get_d(g, "weight"), eg.total_samples, "group"));
            if (g.contains("entities") && !g["entities"].is_none() && py::len(g["entities"]) > 0) {
                // the per-kind partitions, one record per group-local DOF
                // (every DOF covered exactly once)
                std::vector<int32_t> tag(eg.ndof, -1);
                std::vector<int64_t> eoff(eg.ndof, 0);
                std::vector<double> edata;
                for (auto item : g["entities"].cast<py::dict>()) {
                    py::dict ed = item.second.cast<py::dict>();
                    if (!ed.contains("tag"))
                        throw py::value_error("unpack_context: entities carries a group without a tag");
                    const int t = ed["tag"].cast<int>(), kf = ed["kFields"].cast<int>();
                    if (t < 0 || t > 16 || kf != kWireFields[t])
                        throw py::value_error("wire: entity tag " + std::to_string(t) + " (" +
                                              item.first.cast<std::string>() + ") with " + std::to_string(kf) +
                                              " fields is not an entity kind of the wire");
                    IArr loc = get_i(ed, "dof_local");
                    const py::ssize_t k = loc.size();
                    DArr rr = kf ? get_d(ed, "records") : DArr(0);
                    if ((int64_t)rr.size() != (int64_t)k * kf)
                        throw py::value_error("wire: entity records size != count * kFields");
                    std::vector<DArr> sdat;
                    std::vector<py::array_t<int64_t, py::array::c_style | py::array::forcecast>> soff;
                    if (ed.contains("seg") && !ed["seg"].is_none())
                        for (auto sv : ed["seg"].cast<py::list>()) {
                            py::dict sd = sv.cast<py::dict>();
                            sdat.push_back(get_d(sd, "data"));
                            soff.push_back(sd["off"].cast<py::array_t<int64_t, py::array::c_style | py::array::forcecast>>());
                            if (soff.back().size() != k + 1)
                                throw py::value_error("wire: entity seg offsets need count + 1 entries");
                        }
                    for (py::ssize_t i = 0; i < k; ++i) {
                        const int l = loc.data()[i];
                        if (l < 0 || l >= eg.ndof || tag[l] >= 0)
                            throw py::value_error("unpack_context: entity dof_local out of range or repeated");
                        tag[l] = t;
                        if (t == 0) continue;
                        std::vector<Slice> seg;
                        for (size_t q = 0; q < sdat.size(); ++q) {
                            const int64_t a = soff[q].data()[i], b = soff[q].data()[i + 1];
                            if (a < 0 || b < a || b > (int64_t)sdat[q].size())
                                throw py::value_error("wire: entity seg offsets outside their data");
                            seg.push_back({sdat[q].data() + a, b - a});
                        }
                        eoff[l] = (int64_t)edata.size();
                        wire_entity(t, rr.data() + (size_t)i * kf, seg, edata);
                    }
                }
                for (int32_t t : tag)
                    if (t < 0) throw py::value_error("unpack_context: entities do not cover every DOF of the group");
                eg.entity_tag = keep(IArr((py::ssize_t)tag.size(), tag.data()));
                ent_offs.push_back(std::move(eoff));
                ent_datas.push_back(std::move(edata));
                eg.entity_off = ent_offs.back().data();
                eg.entity_data = ent_datas.back().data();
                eg.entity_len = (int64_t)ent_datas.back().size();
            }
            if (g.contains("curv_k") && !g["curv_k"].is_none() && g["curv_k"].cast<int>() > 0) {
                const int K = g["curv_k"].cast<int>();
                IArr cn = get_i(g, "curv_nodes"), cs = get_i(g, "curv_slot");
                DArr cw = get_d(g, "curv_weight");
                if ((int64_t)cn.size() != eg.ndof * K * 4 || (int64_t)cs.size() != eg.ndof * K ||
                    (int64_t)cw.size() != eg.ndof * K)
                    throw py::value_error("validate_context: curvature table sizes != ndof * curv_k (* 4)");
                eg.curv_k = K;
                eg.curv_nodes = keep(cn);
                eg.curv_slot = keep(cs);
                eg.curv_weight = keep(cw);
            }
            if (g.contains("dir_k") && !g["dir_k"].is_none() && g["dir_k"].cast<int>() > 0) {
                // per-DOF directional windows (reference src/bindings.cpp:393-409)
                const int K = g["dir_k"].cast<int>();
                IArr dn = get_i(g, "dir_nodes"), ds = get_i(g, "dir_slot"), dk = get_i(g, "dir_kind");
                DArr dr = get_d(g, "dir_ref"), dw = get_d(g, "dir_weight");
                const int64_t n = eg.ndof * K;
                if ((int64_t)dn.size() != n * 4 || (int64_t)ds.size() != n || (int64_t)dk.size() != n ||
                    (int64_t)dr.size() != n * d || (int64_t)dw.size() != n)
                    throw py::value_error("validate_context: directional table sizes != ndof * dir_k (* 4, * D)");
                eg.dir_k = K;
                eg.dir_eps = g["dir_eps"].cast<double>();
                eg.dir_nodes = keep(dn);
                eg.dir_slot = keep(ds);
                eg.dir_kind = keep(dk);
                eg.dir_ref = keep(dr);
                eg.dir_weight = keep(dw);
                if (g.contains("dir_axis") && !g["dir_axis"].is_none()) eg.dir_axis = keep(get_d(g, "dir_axis"));
            }
            groups.push_back(eg);
        }
        py::dict e = ctx["energy_stencil"].cast<py::dict>();
        es.num_samples = e["num_samples"].cast<int64_t>();
        IArr egc = get_i(e, "gc");
        if ((int64_t)egc.size() != es.num_samples)
            throw py::value_error("validate_context: energy_stencil gc size " + std::to_string(egc.size()) +
                                  " != " + std::to_string(es.num_samples));
        es.gc = keep(egc);
        for (int k = 0; k < d; ++k) {
            const std::string gk = "gn" + std::to_string(k), sk = "s" + std::to_string(k);
            IArr gn = get_i(e, gk.c_str());
            DArr s = get_d(e, sk.c_str());
            if ((int64_t)gn.size() != es.num_samples || (int64_t)s.size() != es.num_samples)
                throw py::value_error("validate_context: energy_stencil " + gk + " size != " +
                                      std::to_string(es.num_samples));
            es.gn[k] = keep(gn);
            es.s[k] = keep(s);
        }
        DArr eW = get_d(e, "W_inv");
        if (eW.size() % dd != 0) throw py::value_error("validate_context: energy_stencil W_inv size");
        es.w_rows = (int64_t)eW.size() / dd;
        es.W_inv = keep(eW);
        if (e.contains("weight") && !e["weight"].is_none())
            es.weight = keep(sample_weights(get_d(e, "weight"), es.num_samples, "energy_stencil"));
        if (e.contains("dir_num") && !e["dir_num"].is_none() && e["dir_num"].cast<int64_t>() > 0) {
            // whole-sample directional table (reference src/bindings.cpp:471-480)
            const int64_t n = e["dir_num"].cast<int64_t>();
            IArr dn = get_i(e, "dir_nodes"), dk = get_i(e, "dir_kind");
            DArr dr = get_d(e, "dir_ref"), dw = get_d(e, "dir_weight");
            if ((int64_t)dn.size() != n * 4 || (int64_t)dk.size() != n || (int64_t)dr.size() != n * d ||
                (int64_t)dw.size() != n)
                throw py::value_error("validate_context: energy-stencil directional table sizes != dir_num (* 4, * D)");
            es.dir_num = n;
            es.dir_eps = e["dir_eps"].cast<double>();
            es.dir_nodes = keep(dn);
            es.dir_kind = keep(dk);
            es.dir_ref = keep(dr);
            es.dir_weight = keep(dw);
            if (e.contains("dir_axis") && !e["dir_axis"].is_none()) es.dir_axis = keep(get_d(e, "dir_axis"));
        }

        // structured tables
        py::list shl = structured["interior_shapes"].cast<py::list>();
        py::list bgl = structured["block_global_dof"].cast<py::list>();
        if (py::len(bgl) != py::len(shl))
            throw py::value_error("structured: block_global_dof count != interior_shapes count");
        for (auto sh : shl) {
            IArr a = sh.cast<IArr>();
            if ((int)a.size() != d) throw py::value_error("structured: each interior shape must have D entries");
            for (int k = 0; k < d; ++k) shapes.push_back(a.data()[k]);
        }
        size_t b = 0;
        for (auto item : bgl) {
            IArr a = item.cast<IArr>();
            int64_t n = 1;
            for (int k = 0; k < d; ++k) n *= shapes[b * d + k];
            if ((int64_t)a.size() != n) throw py::value_error("structured: block_global_dof length != prod(shape)");
            bgd_arrs.push_back(a);
            ++b;
        }
        for (auto& a : bgd_arrs) bgd_ptrs.push_back(a.data());
        hsb = get_i(structured, "halo_src_block");
        hsp = get_i(structured, "halo_src_padded");
        hdb = get_i(structured, "halo_dst_block");
        hdp = get_i(structured, "halo_dst_padded");
        sb = get_i(structured, "sing_block");
        sl = get_i(structured, "sing_logical");
        const int64_t nh = (int64_t)hsb.size();
        if ((int64_t)hdb.size() != nh || (int64_t)hsp.size() != nh * d || (int64_t)hdp.size() != nh * d)
            throw py::value_error("structured: halo table sizes disagree");
        if ((int64_t)sl.size() != (int64_t)sb.size() * d) throw py::value_error("structured: sing table sizes disagree");
        st.num_blocks = (int32_t)bgd_arrs.size();
        st.interior_shapes = shapes.data();
        st.block_global_dof = bgd_ptrs.data();
        st.n_halo = nh;
        st.halo_src_block = hsb.data();
        st.halo_src_padded = hsp.data();
        st.halo_dst_block = hdb.data();
        st.halo_dst_padded = hdp.data();
        st.n_sing = (int64_t)sb.size();
        st.sing_block = sb.data();
        st.sing_logical = sl.data();
    }
};

py::array_t<double> to_array(const std::vector<double>& v)
{
    return py::array_t<double>((py::ssize_t)v.size(), v.data());
}

using FArr = py::array_t<float, py::array::c_style | py::array::forcecast>;

// Learned proposal network (eggmi_mlp_*).
class Mlp
{
  public:
    Mlp(int dim, uint64_t seed, int64_t capacity)
    {
        check(eggmi_mlp_create(require_dim(dim), seed, capacity, &m_));
        check(eggmi_mlp_shape(m_, shape_, &np_, nullptr, nullptr));
    }
    ~Mlp()
    {
        if (m_) eggmi_mlp_destroy(m_);
    }
    Mlp(const Mlp&) = delete;
    Mlp& operator=(const Mlp&) = delete;
    eggmi_mlp* handle() const { return m_; }
    int dim() const { return shape_[0]; }
    py::dict shape() const
    {
        int64_t nd = 0, cap = 0;
        check(eggmi_mlp_shape(m_, nullptr, nullptr, &nd, &cap));
        py::dict d;
        d["dim"] = shape_[0];
        d["n_in"] = shape_[1];
        d["n_in_padded"] = shape_[2];
        d["hidden"] = shape_[3];
        d["n_out"] = shape_[4];
        d["n_params"] = np_;
        d["n_data"] = nd;
        d["capacity"] = cap;
        return d;
    }
    py::array_t<float> get_params() const
    {
        py::array_t<float> out((py::ssize_t)np_);
        check(eggmi_mlp_get_params(m_, out.mutable_data(), np_));
        return out;
    }
    void set_params(const FArr& p)
    {
        if (p.ndim() != 1) throw py::value_error("params must be 1-D");
        check(eggmi_mlp_set_params(m_, p.data(), (int64_t)p.size()));
    }
    int64_t rows_of(const FArr& X, int width, const char* what) const
    {
        if (X.ndim() != 2 || X.shape(1) != width)
            throw py::value_error(std::string(what) + " must have shape (rows, " + std::to_string(width) + ")");
        return (int64_t)X.shape(0);
    }
    py::array_t<float> forward(const FArr& X) const
    {
        const int64_t r = rows_of(X, shape_[2], "X");
        py::array_t<float> Y({(py::ssize_t)r, (py::ssize_t)shape_[4]});
        check(eggmi_mlp_forward(m_, X.data(), r, Y.mutable_data()));
        return Y;
    }
    void data_add(const FArr& X, const FArr& T)
    {
        const int64_t r = rows_of(X, shape_[2], "X");
        if (rows_of(T, shape_[4], "T") != r) throw py::value_error("X and T row counts differ");
        check(eggmi_mlp_data_add(m_, X.data(), T.data(), r));
    }
    py::tuple data_get(int64_t offset, int64_t rows) const
    {
        py::array_t<float> X({(py::ssize_t)rows, (py::ssize_t)shape_[2]});
        py::array_t<float> T({(py::ssize_t)rows, (py::ssize_t)shape_[4]});
        check(eggmi_mlp_data_get(m_, X.mutable_data(), T.mutable_data(), offset, rows));
        return py::make_tuple(X, T);
    }
    void data_clear() { check(eggmi_mlp_data_clear(m_)); }
    py::tuple gradient(const IArr& rows)
    {
        py::array_t<float> g((py::ssize_t)np_);
        double loss = 0.0;
        check(eggmi_mlp_gradient(m_, rows.data(), (int64_t)rows.size(), g.mutable_data(), &loss));
        return py::make_tuple(g, loss);
    }
    py::array_t<double> train(int epochs, int64_t batch, float lr, uint64_t seed)
    {
        py::array_t<double> h((py::ssize_t)std::max(0, epochs));
        check(eggmi_mlp_train(m_, epochs, batch, lr, seed, h.mutable_data()));
        return h;
    }

  private:
    eggmi_mlp* m_ = nullptr;
    int32_t shape_[5] = {};
    int64_t np_ = 0;
};

class Session
{
  public:
    Session(const py::dict& ctx, const py::dict& structured, const DArr& X, const std::string& device, int dim,
            const py::object& control) :
        dim_(require_dim(dim))
    {
        n_ = checked_num_nodes(X, dim_);
        Wire w(ctx, structured, dim_);
        check(eggmi_session_create(dim_, w.groups.data(), (int)w.groups.size(), &w.es, &w.st, X.data(),
                                   (int64_t)n_, device_id(device), &s_));
        if (!control.is_none()) {
            try {
                attach_control(control.cast<py::dict>(), structured);
            } catch (...) {
                eggmi_session_destroy(s_);
                s_ = nullptr;
                throw;
            }
        }
    }

    // The control wire (reference egg/smoothing/control_backend.py:124-136)
    // into eggmi_control; sliding walls are not supported.
    void attach_control(const py::dict& w, const py::dict& structured)
    {
        if (w.contains("blocks")) {
            attach_control_topo(w, structured);
            return;
        }
        for (const char* k : {"walls", "X0"})
            if (w.contains(k) && !w[k].is_none())
                throw py::value_error(std::string("control wire field '") + k +
                                      "': sliding wall controls are not supported by this compute core");
        eggmi_control c {};
        c.block = w.contains("block") ? w["block"].cast<int>() : 0;
        // the block's node shape, for the per-node array lengths
        const py::list shapes = structured["interior_shapes"].cast<py::list>();
        if (c.block < 0 || c.block >= (int)shapes.size())
            throw py::value_error("control wire: block " + std::to_string(c.block) + " out of range");
        IArr bshape = shapes[c.block].cast<IArr>();
        if (bshape.size() != dim_) throw py::value_error("control wire: the block's shape must hold dim values");
        IArr cs = get_i(w, "ctrl_shape");
        if (cs.size() != dim_) throw py::value_error("control wire: ctrl_shape must hold dim values");
        for (int k = 0; k < dim_; ++k) c.ctrl_shape[k] = cs.data()[k];
        c.k_supp = w.contains("k_supp") ? w["k_supp"].cast<int>() : 4;
        std::vector<IArr> offs;
        std::vector<DArr> wts;
        for (int k = 0; k < dim_; ++k) {
            offs.push_back(get_i(w, ("off" + std::to_string(k)).c_str()));
            wts.push_back(get_d(w, ("wt" + std::to_string(k)).c_str()));
            if (offs.back().size() != bshape.data()[k])
                throw py::value_error("control wire: off" + std::to_string(k) + " must hold one value per node of "
                                      "the block's axis " + std::to_string(k));
            if (wts.back().size() != offs.back().size() * c.k_supp)
                throw py::value_error("control wire: wt" + std::to_string(k) + " must hold k_supp values per node");
            c.off[k] = offs.back().data();
            c.wt[k] = wts.back().data();
        }
        DArr C0 = get_d(w, "C0");
        IArr fi = get_i(w, "free_idx");
        int64_t nc = 1;
        for (int k = 0; k < dim_; ++k) nc *= c.ctrl_shape[k];
        if (C0.size() != nc * dim_) throw py::value_error("control wire: C0 must hold dim values per control");
        c.C0 = C0.data();
        c.n_free = fi.size();
        c.free_idx = fi.data();
        DArr b;
        if (w.contains("b") && !w["b"].is_none()) {
            b = get_d(w, "b");
            int64_t nn = dim_;
            for (int k = 0; k < dim_; ++k) nn *= bshape.data()[k];
            if (b.size() != nn) throw py::value_error("control wire: b must hold dim values per node of the block");
            c.b = b.data();
        }
        check(eggmi_session_set_control(s_, &c));
        n_ctrl_ = nc;
        topo_ = false;
    }

    // The multi-block control wire (reference src/bindings.cpp:593-740, the
    // "blocks" key): per-block nets over one stacked control vector, the
    // reduction CSR (nq, r_off, r_col, r_coef) and optional penalty rows
    // (p_off, p_col, p_coef, p_w); sliding-wall frames are not supported.
    void attach_control_topo(const py::dict& w, const py::dict& structured)
    {
        for (const char* k : {"slide", "wall_faces", "seam_face", "arena"})
            if (w.contains(k) && !w[k].is_none())
                throw py::value_error(std::string("control topo wire field '") + k +
                                      "': sliding-wall frames are not supported by this compute core");
        const py::list shapes = structured["interior_shapes"].cast<py::list>();
        const py::list blocks = w["blocks"].cast<py::list>();
        std::vector<eggmi_control> nets;
        std::vector<IArr> offs;
        std::vector<DArr> wts, bs;
        offs.reserve(blocks.size() * dim_);
        wts.reserve(blocks.size() * dim_);
        bs.reserve(blocks.size());
        int64_t n_stack = 0;
        for (const auto item : blocks) {
            const py::dict bd = item.cast<py::dict>();
            eggmi_control c {};
            c.block = bd["block"].cast<int>();
            if (c.block < 0 || c.block >= (int)shapes.size())
                throw py::value_error("control topo: block " + std::to_string(c.block) + " out of range");
            IArr bshape = shapes[c.block].cast<IArr>();
            IArr cs = get_i(bd, "ctrl_shape");
            if (cs.size() != dim_) throw py::value_error("control topo: ctrl_shape must have dim entries");
            int64_t nc = 1, nn = dim_;
            for (int k = 0; k < dim_; ++k) {
                c.ctrl_shape[k] = cs.data()[k];
                nc *= c.ctrl_shape[k];
                nn *= bshape.data()[k];
            }
            c.k_supp = bd["k_supp"].cast<int>();
            for (int k = 0; k < dim_; ++k) {
                offs.push_back(get_i(bd, ("off" + std::to_string(k)).c_str()));
                wts.push_back(get_d(bd, ("wt" + std::to_string(k)).c_str()));
                if (offs.back().size() != bshape.data()[k])
                    throw py::value_error("control topo: off" + std::to_string(k) +
                                          " must hold one value per node of the block's axis " + std::to_string(k));
                if (wts.back().size() != offs.back().size() * c.k_supp)
                    throw py::value_error("control topo: wt" + std::to_string(k) + " size != off size * k_supp");
                c.off[k] = offs.back().data();
                c.wt[k] = wts.back().data();
            }
            bs.push_back(get_d(bd, "b"));
            if (bs.back().size() != nn) throw py::value_error("control topo: b size != prod(node_shape) * dim");
            c.b = bs.back().data();
            n_stack += nc;
            nets.push_back(c);
        }
        DArr C0 = get_d(w, "C0");
        if (C0.size() != n_stack * dim_) throw py::value_error("control topo: C0 size != total controls * dim");
        IArr r_off = get_i(w, "r_off"), r_col = get_i(w, "r_col");
        DArr r_coef = get_d(w, "r_coef");
        if (r_off.size() != n_stack * dim_ + 1 || r_col.size() != r_coef.size() ||
            r_off.data()[r_off.size() - 1] != r_col.size())
            throw py::value_error("control topo: malformed reduction CSR");
        eggmi_control_topo t {};
        t.n_nets = (int32_t)nets.size();
        t.nets = nets.data();
        t.C0 = C0.data();
        t.nq = w["nq"].cast<int64_t>();
        t.r_off = r_off.data();
        t.r_col = r_col.data();
        t.r_coef = r_coef.data();
        IArr p_off, p_col;
        DArr p_coef, p_w;
        if (w.contains("p_off") && !w["p_off"].is_none()) {
            p_off = get_i(w, "p_off");
            p_col = get_i(w, "p_col");
            p_coef = get_d(w, "p_coef");
            p_w = get_d(w, "p_w");
            if (p_col.size() != p_coef.size() ||
                (p_off.size() == 0 ? p_col.size() != 0 : p_off.data()[p_off.size() - 1] != p_col.size()) ||
                (p_off.size() > 0 && p_w.size() != p_off.size() - 1))
                throw py::value_error("control topo: malformed penalty CSR");
            t.n_pen = p_off.size() > 0 ? p_off.size() - 1 : 0;
            t.p_off = p_off.data();
            t.p_col = p_col.data();
            t.p_coef = p_coef.data();
            t.p_w = p_w.data();
        }
        check(eggmi_session_set_control_topo(s_, &t));
        n_ctrl_ = n_stack;
        topo_ = true;
        std::vector<int64_t> nn;
        for (const auto& b : bs) nn.push_back(b.size());
        topo_b_sizes_ = nn;
    }

    void set_control_reduction(int64_t nq, const IArr& off, const IArr& col, const DArr& coef)
    {
        if (!topo_) throw py::value_error("set_control_reduction: session has no multi-block control wire");
        if (off.size() != n_ctrl_ * dim_ + 1 || col.size() != coef.size() || off.data()[off.size() - 1] != col.size())
            throw py::value_error("control topo: malformed reduction CSR");
        check(eggmi_control_set_reduction(s_, nq, off.data(), col.data(), coef.data()));
    }

    void set_control_penalty(const IArr& off, const IArr& col, const DArr& coef, const DArr& w)
    {
        if (!topo_) throw py::value_error("set_control_penalty: session has no multi-block control wire");
        if (col.size() != coef.size() || (off.size() == 0 ? col.size() != 0 : off.data()[off.size() - 1] != col.size()) ||
            (off.size() > 0 && w.size() != off.size() - 1))
            throw py::value_error("control topo: malformed penalty CSR");
        const int64_t n_pen = off.size() > 0 ? off.size() - 1 : 0;
        check(eggmi_control_set_penalty(s_, n_pen, off.data(), col.data(), coef.data(), w.data()));
    }

    void set_control_b(int block_index, const DArr& b)
    {
        if (!topo_) throw py::value_error("set_control_b: session has no multi-block control wire");
        if (block_index < 0 || block_index >= (int)topo_b_sizes_.size())
            throw py::value_error("control topo: set_b block index out of range");
        if (b.size() != topo_b_sizes_[block_index]) throw py::value_error("control topo: set_b size mismatch");
        check(eggmi_control_set_b(s_, block_index, b.data(), (int64_t)b.size()));
    }

    void set_C(const DArr& C) { check(eggmi_control_set_C(s_, C.data(), (int64_t)C.size())); }

    py::array_t<double> get_C()
    {
        if (n_ctrl_ == 0) throw py::value_error("get_C: the session has no control wire");
        py::array_t<double> out(n_ctrl_ * dim_);
        check(eggmi_control_get_C(s_, out.mutable_data(), (int64_t)out.size()));
        return out;
    }

    py::dict run_control(int n_outer, const std::string& phase, double grad_tol, double alpha_min, double lm0,
                         int pcg_max_iter, double pcg_rtol, bool pcg_forcing, bool model_db)
    {
        (void)model_db;  // the Boolean-sum linearization matters for sliding walls only
        if (phase == "untangle")
            throw py::value_error("run_control: the untangle phase stays nodal; run(phase=\"untangle\") first");
        if (n_ctrl_ == 0) throw py::value_error("run_control: the session was built without a control wire");
        eggmi_control_options o {};
        o.n_outer = n_outer;
        o.grad_tol = grad_tol;
        o.alpha_min = alpha_min;
        o.lm0 = lm0;
        o.lm_tries = 8;
        o.pcg_max_iter = pcg_max_iter;
        o.pcg_rtol = pcg_rtol;
        o.pcg_forcing = pcg_forcing ? 1 : 0;
        const int64_t cap = std::max(n_outer, 0) + 1, pcap = cap * o.lm_tries;
        std::vector<double> e(cap), m(cap), a(cap), pi(pcap), pr(pcap);
        int64_t nr = 0, np = 0;
        int32_t iters = 0, conv = 0;
        check(eggmi_run_control(s_, &o, phase_id(phase), e.data(), m.data(), a.data(), cap, &nr, pi.data(), pr.data(),
                                pcap, &np, &iters, &conv));
        e.resize(nr);
        m.resize(nr);
        a.resize(std::max<int64_t>(nr - 1, 0));
        pi.resize(np);
        pr.resize(np);
        py::dict d;
        d["energies"] = to_array(e);
        d["mindets"] = to_array(m);
        d["alphas"] = to_array(a);
        d["frame_jumps"] = to_array(std::vector<double>{});
        d["iters"] = iters;
        d["converged"] = conv != 0;
        d["pcg_iters"] = to_array(pi);
        d["pcg_relres"] = to_array(pr);
        return d;
    }
    ~Session()
    {
        if (s_) eggmi_session_destroy(s_);
    }
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;

    std::tuple<py::array_t<double>, py::array_t<double>> run(int n_sweeps, const std::string& phase, double delta,
                                                             double omega, int report_every)
    {
        const int k = report_every <= 0 ? n_sweeps : report_every;
        const int64_t cap = n_sweeps <= 0 ? 1 : (n_sweeps + k - 1) / k;
        std::vector<double> e(cap), m(cap);
        int64_t n = 0;
        check(eggmi_run(s_, n_sweeps, phase_id(phase), delta, omega, report_every, e.data(), m.data(), cap, &n));
        e.resize(n);
        m.resize(n);
        return {to_array(e), to_array(m)};
    }

    std::tuple<py::array_t<double>, py::array_t<double>> run_fas(int n_cycles, int nu_pre, int nu_post,
                                                                 int nu_coarse, double omega, int max_levels,
                                                                 const std::string& phase)
    {
        if (phase == "untangle")
            throw py::value_error("run_fas: FAS smooths the barrier phases only; use run(phase=\"untangle\") to "
                                  "untangle first");
        const int64_t cap = std::max(1, n_cycles) + 1;
        std::vector<double> e(cap), m(cap);
        int64_t n = 0;
        check(eggmi_run_fas(s_, n_cycles, nu_pre, nu_post, nu_coarse, omega, max_levels, phase_id(phase), e.data(),
                            m.data(), cap, &n));
        e.resize(n);
        m.resize(n);
        return {to_array(e), to_array(m)};
    }

    py::dict mg_level_tables(int level)
    {
        int64_t ns = 0, nh = 0, nsh = 0;
        check(eggmi_mg_level_tables(s_, level, &ns, &nh, &nsh, nullptr, nullptr, nullptr, nullptr));
        py::array_t<int32_t> src(nh + nsh), dst(nh + nsh);
        py::array_t<uint8_t> fr(ns), fi(ns);
        check(eggmi_mg_level_tables(s_, level, &ns, &nh, &nsh, src.mutable_data(), dst.mutable_data(),
                                    fr.mutable_data(), fi.mutable_data()));
        py::dict d;
        d["n_slots"] = ns;
        d["n_halo"] = nh;
        d["n_share"] = nsh;
        d["copy_src"] = src;
        d["copy_dst"] = dst;
        d["free"] = fr;
        d["free_interior"] = fi;
        return d;
    }

    py::array_t<double> mg_probe(int level, int op, const DArr& field)
    {
        int64_t ns = 0;
        check(eggmi_mg_level_tables(s_, level, &ns, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
        int64_t np_ = 0;
        if (level == 1) {
            eggmi_info info {};
            check(eggmi_session_info(s_, &info));
            np_ = info.n_slots;
        } else {
            check(eggmi_mg_level_tables(s_, level - 1, &np_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
        }
        const int64_t n_in = (op == 2 ? ns : np_) * dim_, n_out = (op == 2 ? np_ : ns) * dim_ * (op == 3 ? 2 : 1);
        if (op != 3 && (int64_t)field.size() != n_in)
            throw py::value_error("mg_probe: field must hold dim * n_slots = " + std::to_string(n_in) + " values");
        py::array_t<double> out(n_out);
        check(eggmi_mg_probe(s_, level, op, op == 3 ? nullptr : field.data(), out.mutable_data()));
        return out;
    }

    py::list mg_levels(int max_levels)
    {
        int32_t nl = 0, nb = 0;
        check(eggmi_mg_levels(s_, max_levels, nullptr, 0, &nl, &nb));
        std::vector<int32_t> shp((size_t)nl * nb * dim_ + 1);
        check(eggmi_mg_levels(s_, max_levels, shp.data(), (int64_t)shp.size(), &nl, &nb));
        py::list levels;
        for (int l = 0; l < nl; ++l) {
            py::list blocks;
            for (int b = 0; b < nb; ++b) {
                py::tuple t(dim_);
                for (int k = 0; k < dim_; ++k) t[k] = shp[((size_t)l * nb + b) * dim_ + k];
                blocks.append(t);
            }
            levels.append(blocks);
        }
        return levels;
    }

    py::array_t<double> get_X()
    {
        py::array_t<double> out((py::ssize_t)(n_ * dim_));
        check(eggmi_get_X(s_, out.mutable_data(), (int64_t)(n_ * dim_)));
        return out;
    }

    void set_X(const DArr& X) { check(eggmi_set_X(s_, X.data(), (int64_t)X.size())); }

    py::dict info()
    {
        eggmi_info inf {};
        check(eggmi_session_info(s_, &inf));
        py::dict d;
        d["n_slots"] = inf.n_slots;
        d["n_overflow"] = inf.n_overflow;
        d["n_halo"] = inf.n_halo;
        d["n_fan"] = inf.n_fan;
        d["n_share"] = inf.n_share;
        d["n_tiles"] = inf.n_tiles;
        d["n_stored"] = inf.n_stored;
        d["n_synth"] = inf.n_synth;
        d["structured_energy"] = inf.structured_energy;
        return d;
    }

    py::dict tables()
    {
        eggmi_info inf {};
        check(eggmi_session_info(s_, &inf));
        const int64_t nc = inf.n_halo + inf.n_fan + inf.n_share;
        py::array_t<int32_t> src(nc), dst(nc), g2s((py::ssize_t)n_);
        int32_t nl = 0, nb = 0;
        check(eggmi_mg_levels(s_, 1, nullptr, 0, &nl, &nb));
        py::array_t<int64_t> base(nb);
        check(eggmi_session_tables(s_, src.mutable_data(), dst.mutable_data(), g2s.mutable_data(),
                                   base.mutable_data()));
        py::dict d;
        d["copy_src"] = src;
        d["copy_dst"] = dst;
        d["g2s"] = g2s;
        d["block_base"] = base;
        return d;
    }

    py::array_t<float> packed()
    {
        eggmi_info inf {};
        check(eggmi_session_info(s_, &inf));
        py::array_t<float> out(inf.n_slots * dim_);
        check(eggmi_get_packed(s_, out.mutable_data(), inf.n_slots * dim_));
        return out.attr("reshape")(dim_, inf.n_slots);
    }

    void halo_exchange() { check(eggmi_halo_exchange(s_)); }

    int set_tile_variant(int v)
    {
        int nv = 0;
        check(eggmi_set_tile_variant(s_, v, &nv));
        return nv;
    }

    py::tuple last_timing()
    {
        double ms = 0, kms = 0;
        int64_t n = 0;
        check(eggmi_last_timing(s_, &ms, &n, &kms));
        return py::make_tuple(ms, n, kms);
    }

    int64_t fas_cycle_kernels()
    {
        int64_t n = 0;
        check(eggmi_fas_cycle_kernels(s_, &n));
        return n;
    }

    void set_kernel_timing(bool on) { check(eggmi_set_kernel_timing(s_, on ? 1 : 0)); }
    void set_counters(bool on) { check(eggmi_set_counters(s_, on ? 1 : 0)); }
    void set_graph_chunk(int chunk) { check(eggmi_set_graph_chunk(s_, chunk)); }
    py::array_t<int64_t> last_counters(int n_levels)
    {
        py::array_t<int64_t> out({n_levels, 3});
        check(eggmi_last_counters(s_, out.mutable_data(), n_levels));
        return out;
    }
    py::tuple energy(const std::string& phase, double delta)
    {
        double e = 0.0, m = 0.0;
        check(eggmi_energy(s_, phase_id(phase), delta, &e, &m));
        return py::make_tuple(e, m);
    }
    py::tuple autotune(int n_sweeps, double omega)
    {
        int nv = 0;
        check(eggmi_set_tile_variant(s_, -1, &nv));
        std::vector<double> ms(nv);
        int best = 0;
        check(eggmi_autotune(s_, n_sweeps, omega, ms.data(), nv, &best));
        return py::make_tuple(best, to_array(ms));
    }
    void set_level_variants(const std::vector<int>& v)
    {
        check(eggmi_set_level_variants(s_, v.data(), (int)v.size()));
    }
    int set_multi_sweep(int level, int k, int ty, int tx, int kernel)
    {
        int nk = 0;
        check(eggmi_set_multi_sweep(s_, level, k, ty, tx, kernel, &nk));
        return nk;
    }
    py::dict multi_sweep_info(int level, bool check_tables, unsigned seed)
    {
        int64_t info[8] = {};
        double err = 0.0;
        check(eggmi_multi_sweep_info(s_, level, info, check_tables ? &err : nullptr, seed));
        py::dict d;
        const char* names[8] = {"k", "ty", "tx", "kernel", "tiles", "lds_stride", "core_dofs", "updates"};
        for (int i = 0; i < 8; ++i) d[names[i]] = info[i];
        if (check_tables) d["check_err"] = err;
        return d;
    }
    py::array_t<double> autotune_levels_multi(int n_sweeps, double omega, const std::vector<std::array<int, 4>>& cands,
                                              int max_levels, const std::string& phase)
    {
        std::vector<int32_t> c;
        for (auto& q : cands) c.insert(c.end(), q.begin(), q.end());
        const size_t nc = cands.size() + 1;
        std::vector<double> ms(64 * nc);
        int nl = 0;
        check(eggmi_autotune_levels_multi(s_, n_sweeps, omega, max_levels, phase_id(phase), c.data(),
                                          (int)cands.size(), ms.data(), (int64_t)ms.size(), &nl));
        ms.resize((size_t)nl * nc);
        py::array_t<double> t = to_array(ms);
        return t.attr("reshape")(nl, (int)nc);
    }
    py::array_t<double> autotune_multi(int n_sweeps, double omega, const std::vector<std::array<int, 4>>& cands,
                                       double base_ms)
    {
        std::vector<int32_t> c;
        for (auto& q : cands) c.insert(c.end(), q.begin(), q.end());
        std::vector<double> ms(cands.size());
        check(eggmi_autotune_multi(s_, n_sweeps, omega, c.data(), (int)cands.size(), base_ms, ms.data()));
        return to_array(ms);
    }
    py::tuple autotune_levels(int n_sweeps, double omega, int max_levels, const std::string& phase)
    {
        int nv = 0;
        check(eggmi_set_tile_variant(s_, -1, &nv));
        nv *= 2;  // per-sweep launches, then one-workgroup smoothing
        constexpr int kCap = 64;  // more coarse levels than any hierarchy has
        std::vector<double> ms((size_t)kCap * nv);
        std::vector<int> best(kCap);
        int nl = 0;
        check(eggmi_autotune_levels(s_, n_sweeps, omega, max_levels, phase_id(phase), ms.data(),
                                    (int64_t)ms.size(), best.data(), kCap, &nl));
        ms.resize((size_t)nl * nv);
        best.resize(nl);
        py::array_t<double> t = to_array(ms);
        return py::make_tuple(py::cast(best), t.attr("reshape")(nl, nv));
    }
    int64_t harvest(Mlp& m, int n_snap, int every, int horizon, double omega)
    {
        int64_t added = 0;
        check(eggmi_session_harvest(s_, m.handle(), n_snap, every, horizon, omega, &added));
        return added;
    }
    void set_proposal(const py::object& m, const std::string& mode, float scale, int every)
    {
        int md;
        if (mode == "replace")
            md = 0;
        else if (mode == "predict_correct")
            md = 1;
        else
            throw py::value_error("proposal mode must be 'replace' or 'predict_correct'; got '" + mode + "'");
        if (m.is_none()) {
            check(eggmi_session_set_proposal(s_, nullptr, md, scale));
            proposal_ = py::none();
            return;
        }
        Mlp& net = m.cast<Mlp&>();
        check(eggmi_session_set_proposal_every(s_, every));
        check(eggmi_session_set_proposal(s_, net.handle(), md, scale));
        proposal_ = m;  // the session borrows the network: keep it alive
    }

  private:
    int dim_;
    size_t n_ = 0;
    int64_t n_ctrl_ = 0;
    bool topo_ = false;                   // a multi-block control wire is attached
    std::vector<int64_t> topo_b_sizes_;   // its nets' b lengths
    eggmi_session* s_ = nullptr;
    py::object proposal_ = py::none();
};

}  // namespace

double& last_pcg_ms()
{
    static double ms = 0.0;
    return ms;
}

PYBIND11_MODULE(cpp_core, m)
{
    m.doc() = "egg compute core for AMD Instinct MI355X (HIP, gfx950) over the eggmi C ABI.";
    m.def("ping", [] { return eggmi_ping(); }, "Liveness check; returns 0.");
    // FP32 bulk arithmetic (FP64 reductions); tolerance floors follow the
    // reference's float build.
    m.attr("REAL_IS_FLOAT") = true;

    py::class_<Session>(m, "CppStructuredSweepSession",
                        "Persistent device-resident structured smoothing session on the MI355X.")
      .def(py::init<const py::dict&, const py::dict&, const DArr&, const std::string&, int, const py::object&>(),
           py::arg("ctx_arrays"), py::arg("structured"), py::arg("X"), py::kw_only(), py::arg("device") = "auto",
           py::arg("dim") = 2, py::arg("control") = py::none())
      .def("run", &Session::run, py::arg("n_sweeps"), py::kw_only(), py::arg("phase") = "barrier",
           py::arg("delta") = 0.0, py::arg("omega") = 1.0, py::arg("report_every") = 0,
           "Run n_sweeps of block-Jacobi on the resident store; returns (energies, mindets).")
      .def("run_fas", &Session::run_fas, py::arg("n_cycles"), py::kw_only(), py::arg("nu_pre") = 2,
           py::arg("nu_post") = 2, py::arg("nu_coarse") = 32, py::arg("omega") = 1.0, py::arg("max_levels") = 32,
           py::arg("phase") = "barrier", "Run n_cycles FAS V-cycles; returns per-cycle (energies, mindets).")
      .def("mg_probe", &Session::mg_probe, py::arg("level"), py::arg("op"), py::arg("field"),
           "FAS transfer operator `op` (0 inject, 1 full weighting, 2 prolongation, 3 tau check) between "
           "coarse level `level` and its parent on a [dim][n_slots] field.")
      .def("mg_level_tables", &Session::mg_level_tables, py::arg("level"),
           "Copy tables (halo | share, node slots) and relaxed-DOF masks of FAS coarse level `level` (>= 1).")
      .def("mg_levels", &Session::mg_levels, py::arg("max_levels") = 32,
           "Coarse-hierarchy shape: one list per coarse level of per-block interior node-count tuples.")
      .def("get_X", &Session::get_X, "Download X gathered back to global node order, shape (N*dim,).")
      .def("set_X", &Session::set_X, py::arg("X"), "Upload X (global order, flat) into the resident store.")
      .def("info", &Session::info)
      .def("tables", &Session::tables)
      .def("packed", &Session::packed)
      .def("halo_exchange", &Session::halo_exchange)
      .def("set_tile_variant", &Session::set_tile_variant, py::arg("variant"))
      .def("last_timing", &Session::last_timing)
      .def("set_C", &Session::set_C, py::arg("C"),
           "Upload a full control lattice (n_ctrl * dim values) and re-evaluate the net into the state.")
      .def("get_C", &Session::get_C, "Download the full control lattice (n_ctrl * dim values).")
      .def("set_control_reduction", &Session::set_control_reduction, py::arg("nq"), py::arg("off"), py::arg("col"),
           py::arg("coef"), "Replace the multi-block reduction CSR dC = R dq (same stacked-component rows).")
      .def("set_control_penalty", &Session::set_control_penalty, py::arg("off"), py::arg("col"), py::arg("coef"),
           py::arg("w"), "Replace the multi-block penalty rows (energy 1/2 sum w (P C)^2).")
      .def("set_control_b", &Session::set_control_b, py::arg("block_index"), py::arg("b"),
           "Replace one net's Boolean-sum offset and re-evaluate the nets into the state.")
      .def("run_control", &Session::run_control, py::arg("n_outer"), py::kw_only(), py::arg("phase") = "barrier",
           py::arg("grad_tol") = 1e-8, py::arg("alpha_min") = 1e-4, py::arg("lm0") = 0.0,
           py::arg("pcg_max_iter") = 200, py::arg("pcg_rtol") = 1e-10, py::arg("pcg_forcing") = true,
           py::arg("model_db") = true,
           "Up to n_outer control-net reduced Gauss-Newton iterations on the resident net; returns {energies, "
           "mindets, alphas, frame_jumps, iters, converged, pcg_iters, pcg_relres}.")
      .def("fas_cycle_kernels", &Session::fas_cycle_kernels)
      .def("set_kernel_timing", &Session::set_kernel_timing, py::arg("on"))
      .def("set_counters", &Session::set_counters, py::arg("on"))
      .def("set_graph_chunk", &Session::set_graph_chunk, py::arg("chunk"),
           "Replay plain sweeps as captured hipGraph chunks of `chunk` sweeps (even; 0 = off), and FAS "
           "cycles as captured graphs (chunk > 0).")
      .def("last_counters", &Session::last_counters, py::arg("n_levels") = 1,
           "Counters of the last run (counting on): rows per level of (line-search trials, DOF updates, "
           "accepted learned proposals).")
      .def("energy", &Session::energy, py::kw_only(), py::arg("phase") = "barrier", py::arg("delta") = 0.0,
           "(energy, min det) of the resident state, reduced on the device in FP64.")
      .def("autotune", &Session::autotune, py::arg("n_sweeps") = 20, py::arg("omega") = 0.8,
           "Time every tile variant on the current state (restored), keep the fastest; returns (best, ms).")
      .def("set_multi_sweep", &Session::set_multi_sweep, py::arg("level"), py::arg("k"), py::arg("ty") = 16,
           py::arg("tx") = 16, py::arg("kernel") = 0,
           "k block-Jacobi sweeps per launch on a level (0 = fine, i = FAS coarse level i) over ty x tx core "
           "tiles with LDS-staged dependency closures (the results of k launches up to a few float ulps); k <= 1 turns it off. Returns the "
           "number of multi-sweep kernel variants.")
      .def("multi_sweep_info", &Session::multi_sweep_info, py::arg("level") = 0, py::arg("check") = false,
           py::arg("seed") = 1,
           "A level's multi-sweep tables (k, tile, kernel, tiles, LDS stride, core DOFs, DOF updates per launch); "
           "check=True adds check_err, the host-side schedule check of the tables (0 when right).")
      .def("autotune_multi", &Session::autotune_multi, py::arg("n_sweeps"), py::arg("omega"), py::arg("cands"),
           py::arg("base_ms") = std::numeric_limits<double>::infinity(),
           "Time multi-sweep candidates (k, ty, tx, kernel) on the fine level from the current state (restored) "
           "and keep the fastest if it beats base_ms; returns ms per candidate (inf: not buildable).")
      .def("autotune_levels_multi", &Session::autotune_levels_multi, py::arg("n_sweeps"), py::arg("omega"),
           py::arg("cands"), py::arg("max_levels") = 32, py::arg("phase") = "barrier",
           "Per FAS coarse level, time one smoothing call with the level's one-sweep setting and per multi-sweep "
           "candidate (k, ty, tx, kernel), keep the fastest; returns ms [level, 1 + candidates] (-1: not "
           "buildable).")
      .def("set_level_variants", &Session::set_level_variants, py::arg("variants"),
           "Tile variant of each FAS coarse level (variants[i] for level i + 1; -1 = the session's); V + v "
           "(V = the tile variant count) smooths the level in one single-workgroup launch on variant v's tiles.")
      .def("autotune_levels", &Session::autotune_levels, py::arg("n_sweeps") = 16, py::arg("omega") = 0.8,
           py::arg("max_levels") = 32, py::arg("phase") = "barrier",
           "Time every tile variant on each FAS coarse level from the restricted resident state, as per-sweep "
           "launches and (small levels) as one single-workgroup launch, and keep each level's fastest (the fine "
           "state is unchanged); returns (best per level, ms [level, 2 V]; -1 = not timed).")
      .def("harvest", &Session::harvest, py::arg("mlp"), py::arg("n_snap"), py::kw_only(), py::arg("every") = 8,
           py::arg("horizon") = 8, py::arg("omega") = 0.8,
           "Append (features, displacement over `horizon` plain sweeps) pairs of every interior DOF to the "
           "network's training set; returns the number of pairs added.")
      .def("set_proposal", &Session::set_proposal, py::arg("mlp"), py::kw_only(), py::arg("mode") = "replace",
           py::arg("scale") = 1.0f, py::arg("every") = 1,
           "Use a trained network (or None) as the learned proposal of subsequent run() calls, in sweeps 0, every, "
           "2 every, ... of each call (FAS fine smoothing: every sweep, or with every > 1 each call's first).");

    py::class_<Mlp>(m, "ProposalMLP", "Learned Newton-step proposal network (BF16 matrix cores, FP32 master).")
      .def(py::init<int, uint64_t, int64_t>(), py::arg("dim") = 2, py::arg("seed") = 0,
           py::arg("capacity") = 0)
      .def_property_readonly("dim", &Mlp::dim)
      .def("shape", &Mlp::shape)
      .def("get_params", &Mlp::get_params)
      .def("set_params", &Mlp::set_params, py::arg("params"))
      .def("forward", &Mlp::forward, py::arg("X"))
      .def("data_add", &Mlp::data_add, py::arg("X"), py::arg("T"))
      .def("data_get", &Mlp::data_get, py::arg("offset"), py::arg("rows"))
      .def("data_clear", &Mlp::data_clear)
      .def("gradient", &Mlp::gradient, py::arg("rows"))
      .def("train", &Mlp::train, py::arg("epochs"), py::kw_only(), py::arg("batch") = 4096,
           py::arg("lr") = 1e-3, py::arg("seed") = 0);

    m.def(
      "directional_eval",
      [](const py::array_t<double, py::array::c_style | py::array::forcecast>& X, const IArr& nodes,
         const IArr& kind, const py::array_t<double, py::array::c_style | py::array::forcecast>& ref,
         const DArr& weight, double eps) {
          // host-side total and analytic gradient of a directional sample
          // table (reference src/bindings.cpp:296-330, :1819-1870), in FP64
          if (X.ndim() != 2 || (X.shape(1) != 2 && X.shape(1) != 3))
              throw py::value_error("directional_eval: X must be (n_nodes, 2) or (n_nodes, 3)");
          const int d = (int)X.shape(1);
          const py::ssize_t n = weight.size(), nn = X.shape(0);
          if (nodes.ndim() != 2 || nodes.shape(1) != 4 || nodes.shape(0) != n || kind.size() != n ||
              ref.ndim() != 2 || ref.shape(0) != n || ref.shape(1) != d)
              throw py::value_error("directional_eval: inconsistent sample-table shapes");
          const double* x = X.data();
          py::array_t<double> grad({nn, (py::ssize_t)d});
          double* g = grad.mutable_data();
          std::fill(g, g + nn * d, 0.0);
          double e = 0.0;
          for (py::ssize_t i = 0; i < n; ++i) {
              const int k = kind.data()[i];
              const int ar = k == 1 ? 3 : (k == 0 || k == 2) ? 2 : 0;
              if (!ar) throw py::value_error("directional_eval: unsupported sample kind " + std::to_string(k));
              const int* nd = nodes.data() + 4 * i;
              for (int m = 0; m < ar; ++m)
                  if (nd[m] < 0 || nd[m] >= nn) throw py::value_error("directional_eval: node id out of range");
              const double* r = ref.data() + (py::ssize_t)d * i;
              const double w = weight.data()[i];
              auto P = [&](int m, int c) { return x[(py::ssize_t)nd[m] * d + c]; };
              if (ar == 2) {  // parallel: u = p1 - p0; stem: u = p0 - p1
                  const int hd = k == 0 ? 1 : 0;
                  double u[3], ss = eps * eps, dot = 0.0;
                  for (int c = 0; c < d; ++c) {
                      u[c] = P(hd, c) - P(1 - hd, c);
                      ss += u[c] * u[c];
                      dot += u[c] * r[c];
                  }
                  const double sn = std::sqrt(ss), cc = dot / sn;
                  e += w * cc * cc;
                  for (int c = 0; c < d; ++c) {
                      const double gu = 2.0 * w * cc * (r[c] / sn - cc * u[c] / ss);
                      g[(py::ssize_t)nd[hd] * d + c] += gu;
                      g[(py::ssize_t)nd[1 - hd] * d + c] -= gu;
                  }
              } else {  // line (a, c, b): r = q1/s1 + q2/s2, qk away from c
                  double q[2][3], s[2], rv[3], qr[2] = {0.0, 0.0}, rr = 0.0;
                  for (int t = 0; t < 2; ++t) {
                      double ss = eps * eps;
                      for (int c = 0; c < d; ++c) {
                          q[t][c] = P(2 * t, c) - P(1, c);
                          ss += q[t][c] * q[t][c];
                      }
                      s[t] = std::sqrt(ss);
                  }
                  for (int c = 0; c < d; ++c) {
                      rv[c] = q[0][c] / s[0] + q[1][c] / s[1];
                      rr += rv[c] * rv[c];
                  }
                  e += w * rr;
                  for (int t = 0; t < 2; ++t)
                      for (int c = 0; c < d; ++c) qr[t] += q[t][c] * rv[c];
                  for (int c = 0; c < d; ++c) {
                      double gs = 0.0;
                      for (int t = 0; t < 2; ++t) {
                          const double gt = 2.0 * w / s[t] * (rv[c] - q[t][c] * qr[t] / (s[t] * s[t]));
                          g[(py::ssize_t)nd[2 * t] * d + c] += gt;
                          gs += gt;
                      }
                      g[(py::ssize_t)nd[1] * d + c] -= gs;
                  }
              }
          }
          return py::make_tuple(e, grad);
      },
      py::arg("X"), py::arg("nodes"), py::arg("kind"), py::arg("ref"), py::arg("weight"), py::arg("eps") = 1e-12,
      "Directional soft-energy total and analytic gradient over a sample table (host side, FP64): X (n_nodes, d), "
      "nodes (n, 4) with -1 pads, kind (n,) 0 parallel / 1 line / 2 stem, ref (n, d), weight (n,).");

    // geometry entity calls (reference src/bindings.cpp:1872-1990): the
    // single-point calls take the positional record (2D fixed-size kinds; 3D
    // kinds and tag 0 are free there, as in the reference's 2D dispatch),
    // the batch call one entity's SoA record and payloads
    auto positional = [](const DArr& params, int tag) {
        std::vector<double> b {double(tag)};
        if (tag == 10 || tag == 11)
            throw py::value_error("geometry: tag " + std::to_string(tag) +
                                  " keeps its data in an arena; use geometry_project_batch");
        if (tag < 0 || tag > 16) throw py::value_error("geometry: unknown entity tag " + std::to_string(tag));
        const int nf = kWireFields[tag];
        if ((int)params.size() < nf)
            throw py::value_error("geometry: tag " + std::to_string(tag) + " needs " + std::to_string(nf) +
                                  " parameters");
        b.insert(b.end(), params.data(), params.data() + nf);
        return b;
    };
    auto point_frame = [positional](const DArr& p, int tag, const DArr& params, bool foot) {
        if (p.size() < 2) throw py::value_error("geometry: p must hold 2 coordinates");
        const bool free2 = tag == 0 || tag == 4 || tag == 5 || tag == 12 || tag == 13 || tag == 14 || tag == 16;
        const std::vector<double> b = free2 ? std::vector<double> {0.0} : positional(params, tag);
        py::array_t<double> out(2);
        const int K = free2 ? 2 : 1;
        std::vector<double> f(2), B(2 * K);
        check(eggmi_geometry_eval(2, b.data(), (int64_t)b.size(), 1, p.data(), f.data(), B.data()));
        out.mutable_data()[0] = foot ? f[0] : B[0];
        out.mutable_data()[1] = foot ? f[1] : B[K];
        return out;
    };
    m.def(
      "geometry_project", [point_frame](const DArr& p, int tag, const DArr& params) {
          return point_frame(p, tag, params, true);
      },
      py::arg("p"), py::arg("tag"), py::arg("params"), "Project point onto entity. Returns proj(2,).");
    m.def(
      "geometry_tangent", [point_frame](const DArr& p, int tag, const DArr& params) {
          return point_frame(p, tag, params, false);
      },
      py::arg("p"), py::arg("tag"), py::arg("params"), "Tangent direction at point on entity. Returns tang(2,).");
    m.def(
      "geometry_project_batch",
      [](const DArr& q, int tag, const DArr& records, const py::list& seg) {
          if (q.ndim() != 2 || (q.shape(1) != 2 && q.shape(1) != 3))
              throw py::value_error("geometry_project_batch: q must be (n, 2) or (n, 3)");
          if (tag == 0) throw py::value_error("geometry_project_batch: tag 0 (free) has no foot");
          if (tag < 0 || tag > 16) throw py::value_error("geometry_project_batch: unknown entity tag");
          const int dim = (int)q.shape(1);
          if ((int)records.size() < kWireFields[tag])
              throw py::value_error("geometry_project_batch: record shorter than the kind's fields");
          std::vector<DArr> keep;
          std::vector<Slice> slices;
          for (auto item : seg) {
              py::dict sd = item.cast<py::dict>();
              keep.push_back(get_d(sd, "data"));
              auto off = sd["off"].cast<py::array_t<int64_t, py::array::c_style | py::array::forcecast>>();
              if (off.size() < 2 || off.data()[0] < 0 || off.data()[1] < off.data()[0] ||
                  off.data()[1] > (int64_t)keep.back().size())
                  throw py::value_error("geometry_project_batch: seg offsets outside their data");
              slices.push_back({keep.back().data() + off.data()[0], off.data()[1] - off.data()[0]});
          }
          std::vector<double> b;
          wire_entity(tag, records.data(), slices, b);
          const py::ssize_t n = q.shape(0);
          const int K = (tag == 4 || tag == 5 || tag == 12 || tag == 14) ? 2 : 1;
          py::array_t<double> feet({n, (py::ssize_t)dim}), basis({n, (py::ssize_t)dim, (py::ssize_t)K});
          check(eggmi_geometry_eval(dim, b.data(), (int64_t)b.size(), n, q.data(), feet.mutable_data(),
                                    basis.mutable_data()));
          return py::make_tuple(feet, basis);
      },
      py::arg("q"), py::arg("tag"), py::arg("records"), py::arg("seg"),
      "Project (n, d) points onto one entity given its SoA record and payloads: (feet (n, d), orthonormal tangent "
      "basis (n, d, tdim)).");

    m.def(
      "pcg_csr",
      [](const py::array_t<int64_t, py::array::c_style | py::array::forcecast>& rowptr, const IArr& col,
         const DArr& val, const DArr& b, double rtol, int max_iter, const std::string& recurrence) {
          const int64_t n = (int64_t)b.size();
          if ((int64_t)rowptr.size() != n + 1) throw py::value_error("pcg: rowptr must have n+1 entries");
          if (col.size() != val.size()) throw py::value_error("pcg: col and val sizes differ");
          if (n > 0 && rowptr.data()[n] != (int64_t)col.size()) throw py::value_error("pcg: rowptr[n] != nnz");
          py::array_t<double> x(n);
          int iters = 0, status = 0;
          double relres = 0.0, ms = 0.0;
          check(eggmi_pcg_csr(n, rowptr.data(), col.data(), val.data(), b.data(), x.mutable_data(), rtol, max_iter,
                              recurrence_id(recurrence), &iters, &relres, &status, &ms));
          last_pcg_ms() = ms;
          return py::make_tuple(x, iters, relres, status);
      },
      py::arg("rowptr"), py::arg("col"), py::arg("val"), py::arg("b"), py::kw_only(), py::arg("rtol") = 1e-10,
      py::arg("max_iter") = 200, py::arg("recurrence") = "standard",
      "Jacobi-preconditioned CG (FP64, on the GPU) for a CSR operator; recurrence 'standard' (the reference's) "
      "or 'one_reduction' (Chronopoulos-Gear); returns (x, iters, relres, status).");
    m.def(
      "pcg_lattice",
      [](const IArr& shape, double diag, double off, const DArr& b, double rtol, int max_iter,
         const std::string& recurrence) {
          const int dim = (int)shape.size();
          py::array_t<double> x((py::ssize_t)b.size());
          int64_t n = 1;
          for (int k = 0; k < dim; ++k) n *= shape.data()[k];
          if (n != (int64_t)b.size()) throw py::value_error("pcg: b must hold prod(shape) entries");
          int iters = 0, status = 0;
          double relres = 0.0, ms = 0.0;
          check(eggmi_pcg_lattice(dim, shape.data(), diag, off, b.data(), x.mutable_data(), rtol, max_iter,
                                  recurrence_id(recurrence), &iters, &relres, &status, &ms));
          last_pcg_ms() = ms;
          return py::make_tuple(x, iters, relres, status);
      },
      py::arg("shape"), py::arg("diag"), py::arg("off"), py::arg("b"), py::kw_only(), py::arg("rtol") = 1e-10,
      py::arg("max_iter") = 200, py::arg("recurrence") = "standard",
      "Jacobi-preconditioned CG (FP64, on the GPU) for the matrix-free (2 dim + 1)-point lattice operator "
      "(diag, off; Dirichlet closure); returns (x, iters, relres, status).");
    m.def("last_pcg_ms", [] { return last_pcg_ms(); },
          "Device time (ms) of the last pcg_csr / pcg_lattice solve, uploads excluded.");

    m.def(
      "cpp_structured_sweep",
      [](const py::dict& ctx, const py::dict& structured, const DArr& X, int n_sweeps, const std::string& device,
         const std::string& phase, double delta, int dim, double omega, int report_every) {
          Session s(ctx, structured, X, device, dim, py::none());
          auto [e, m] = s.run(n_sweeps, phase, delta, omega, report_every);
          return py::make_tuple(s.get_X(), e, m);
      },
      py::arg("ctx_arrays"), py::arg("structured"), py::arg("X"), py::arg("n_sweeps"), py::kw_only(),
      py::arg("device") = "auto", py::arg("phase") = "barrier", py::arg("delta") = 0.0, py::arg("dim") = 2,
      py::arg("omega") = 1.0, py::arg("report_every") = 0,
      "Run n_sweeps of block-Jacobi over a halo-padded structured store; returns (X_out, energies, mindets).");
}
