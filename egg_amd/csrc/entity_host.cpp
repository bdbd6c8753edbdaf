// Host-side checks of entity blobs (layout in entity.hpp) before they reach
// the device: every count that sizes a payload, every index a projection
// computes (knot spans, control points, trim loops, composite segments) and
// every value the kernels read must lie inside the blob. The caps and the
// rejection of nested composites follow the reference (src/geometry.hpp:
// 686-722, :965-990).
#include "entity.hpp"
#include "session.hpp"

#include <cmath>
#include <string>

namespace eggmi
{
namespace
{

[[noreturn]] void bad(const std::string& what) { throw InvalidArgument("entity: " + what); }

// an exact non-negative integer stored as a double
long long as_count(double v, const char* what)
{
    if (!std::isfinite(v) || v < 0.0 || v != std::floor(v) || v > 1e9) bad(std::string(what) + " must be a count");
    return (long long)v;
}

void need(long long have, long long want, const char* kind)
{
    if (want > have) bad(std::string(kind) + " record is shorter than its sizes say");
}

void finite(const double* v, long long n, const char* kind)
{
    for (long long i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) bad(std::string(kind) + " record holds a non-finite value");
}

void knots_ok(const double* U, long long len, const char* kind)
{
    for (long long i = 1; i < len; ++i)
        if (U[i] < U[i - 1]) bad(std::string(kind) + " knots must be non-decreasing");
}

long long check(const double* b, long long avail, int dim, bool in_composite)
{
    if (avail < 1) bad("empty record");
    const double tv = b[0];
    if (!std::isfinite(tv) || tv != std::floor(tv) || tv < 1 || tv > (double)kEntMaxTag || tv == (double)kEntFree)
        bad("tag " + std::to_string(tv) + " is not an entity kind");
    const int tag = (int)tv;
    if (ent_dim(tag) != dim)
        bad("tag " + std::to_string(tag) + " is a " + std::to_string(ent_dim(tag)) + "D entity in a " +
            std::to_string(dim) + "D session");
    if (in_composite && !(tag == kEntSegment || tag == kEntCircle || tag == kEntEllipse || tag == kEntRail ||
                          (tag >= kEntCircleArc && tag <= kEntBSpline)))
        bad("a composite segment must be a 2D curve (no nested composite)");
    const double* f = b + 1;
    const long long have = avail - 1;
    auto fixed = [&](long long n, const char* kind) {
        need(have, n, kind);
        finite(f, n, kind);
        return 1 + n;
    };
    switch (tag) {
    case kEntSegment:
    case kEntRail: return fixed(4, "segment");
    case kEntCircle: return fixed(3, "circle");
    case kEntEllipse: {
        const long long n = fixed(4, "ellipse");
        if (f[2] == 0.0 || f[3] == 0.0) bad("ellipse radii must be nonzero");
        return n;
    }
    case kEntCircleArc: return fixed(6, "circle arc");
    case kEntEllipseArc: return fixed(8, "ellipse arc");
    case kEntQuadBezier: return fixed(8, "quadratic Bezier");
    case kEntCubicBezier: return fixed(10, "cubic Bezier");
    case kEntSphere: return fixed(13, "sphere");
    case kEntPlane: return fixed(9, "plane");
    case kEntCylinder: return fixed(13, "cylinder");
    case kEntLine3: return fixed(8, "3D line");
    case kEntRail3: return fixed(6, "3D rail");
    case kEntBSpline: {
        need(have, 6, "B-spline");
        finite(f, 6, "B-spline");
        const long long p = as_count(f[0], "B-spline degree"), n = as_count(f[1], "B-spline control count");
        if (p < 1 || p > kEntCurveCap - 1)
            bad("B-spline curve degree " + std::to_string(p) + " outside 1.." + std::to_string(kEntCurveCap - 1));
        if (n < p + 1) bad("a B-spline needs at least degree + 1 control points");
        const bool rat = f[5] != 0.0;
        if (f[5] != 0.0 && f[5] != 1.0) bad("B-spline has_w must be 0 or 1");
        const long long nk = n + p + 1, len = 6 + nk + 2 * n + (rat ? n : 0);
        need(have, len, "B-spline");
        finite(f, len, "B-spline");
        knots_ok(f + 6, nk, "B-spline");
        if (!(f[6 + n] > f[6 + p])) bad("B-spline live knot domain is empty");
        if (rat)
            for (long long i = 0; i < n; ++i)
                if (!(f[6 + nk + 2 * n + i] > 0.0)) bad("NURBS weights must be positive");
        return 1 + len;
    }
    case kEntBSplineSurf: {
        need(have, 6, "B-spline surface");
        finite(f, 6, "B-spline surface");
        const long long pu = as_count(f[0], "surface degree"), pv = as_count(f[1], "surface degree");
        const long long nu = as_count(f[2], "surface control count"), nv = as_count(f[3], "surface control count");
        if (pu < 1 || pv < 1 || pu > kEntSurfCap - 1 || pv > kEntSurfCap - 1)
            bad("B-spline surface degrees outside 1.." + std::to_string(kEntSurfCap - 1));
        if (nu < pu + 1 || nv < pv + 1) bad("a B-spline surface needs degree + 1 control points per direction");
        if (f[4] != 0.0 && f[4] != 1.0) bad("B-spline surface has_w must be 0 or 1");
        const bool rat = f[4] != 0.0;
        const long long nl = as_count(f[5], "trim loop table size");
        const long long nku = nu + pu + 1, nkv = nv + pv + 1, net = nu * nv;
        long long len = 6 + nku + nkv + 3 * net + (rat ? net : 0) + nl;
        need(have, len, "B-spline surface");
        finite(f, len, "B-spline surface");
        knots_ok(f + 6, nku, "B-spline surface (u)");
        knots_ok(f + 6 + nku, nkv, "B-spline surface (v)");
        if (!(f[6 + nu] > f[6 + pu]) || !(f[6 + nku + nv] > f[6 + nku + pv]))
            bad("B-spline surface live knot domain is empty");
        if (rat)
            for (long long i = 0; i < net; ++i)
                if (!(f[6 + nku + nkv + 3 * net + i] > 0.0)) bad("NURBS weights must be positive");
        long long nverts = 0;
        if (nl >= 2) {
            const double* loops = f + len - nl;
            if (loops[0] != 0.0) bad("trim loop table must start at 0");
            long long prev = 0;
            for (long long i = 1; i < nl; ++i) {
                const long long v = as_count(loops[i], "trim loop offset");
                if (v < prev) bad("trim loop offsets must be non-decreasing");
                prev = v;
            }
            nverts = prev;
        }
        need(have, len + 2 * nverts, "B-spline surface");
        finite(f + len, 2 * nverts, "B-spline surface");
        return 1 + len + 2 * nverts;
    }
    case kEntComposite: {
        if (in_composite) bad("nested composite");
        need(have, 1, "composite");
        const long long m = as_count(f[0], "composite segment count");
        if (m < 1) bad("a composite needs at least one segment");
        need(have, 1 + m, "composite");
        long long end = 2 + m;  // blob values used so far
        for (long long i = 0; i < m; ++i) {
            const long long off = as_count(f[1 + i], "composite segment offset");
            if (off < 2 + m || off >= avail) bad("composite segment offset outside its record");
            const long long len = check(b + off, avail - off, dim, true);
            end = std::max(end, off + len);
        }
        return end;
    }
    default: bad("tag " + std::to_string(tag) + " is not an entity kind");
    }
}

}  // namespace

long long entity_blob_check(const double* b, long long avail, int dim)
{
    const long long len = check(b, avail, dim, false);
    // the device reads the record in FP32: every count and offset in it must
    // stay an exact float
    if (len >= (1LL << 24)) bad("a record of 2^24 values or more");
    return len;
}

}  // namespace eggmi
