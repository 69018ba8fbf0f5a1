// VTK frame writer: the legacy-VTK ASCII POLYDATA frame of the reference's
// vtkOutputUnif (ParticlesEdgesPanels2.f90:2116-2260), written natively and
// byte for byte as the reference's Fortran runtime writes it.
//
// Host code only.  The large blocks (point coordinates, per-point and
// per-panel scalars) are formatted by several threads into separate buffers
// and written in order.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "bve_host_error.h"
#include "../../include/bve.h"

namespace bve {
std::string &context_free_error()
{
    thread_local std::string message;
    return message;
}
}  // namespace bve

namespace {

constexpr int kTriPanel = 3, kQuadPanel = 4, kVoronoiPanel = 10;  // ParticlesEdgesPanels2.f90:36-38

// Fortran F24.18: the value rounded to 18 decimals (to nearest, ties to
// even), a "0" before the point when the integer part is zero, "-" for
// every value with the sign bit set (also -0 and negatives that round to
// zero), right justified in 24 columns; 24 '*' when it does not fit; NaN,
// Inf and -Inf right justified.  The rounding is exact: |v| = m 2^e, so
// |v| 10^18 = m 5^18 2^(e+18), an integer product of at most 95 bits
// shifted right with round-half-even.
void put_f24_18(std::string &out, double v)
{
    constexpr int kWidth = 24;
    char buf[64];
    int len = 0;
    if (std::isnan(v)) {
        len = snprintf(buf, sizeof buf, "NaN");
    } else if (std::isinf(v)) {
        len = snprintf(buf, sizeof buf, v > 0 ? "Inf" : "-Inf");
    } else {
        const bool neg = std::signbit(v);
        const double a = std::fabs(v);
        if (a >= 131072.0) {  // at least six integer digits: never fits
            out.append(kWidth, '*');
            return;
        }
        uint64_t bits;
        std::memcpy(&bits, &a, sizeof bits);
        const int biased = int((bits >> 52) & 0x7ff);
        uint64_t m = bits & ((uint64_t(1) << 52) - 1);
        int e;
        if (biased == 0) {
            e = -1074;
        } else {
            m |= uint64_t(1) << 52;
            e = biased - 1075;
        }
        // a < 2^17 and m >= 2^52 (or a subnormal) give e + 18 < 0
        const unsigned __int128 p = (unsigned __int128)m * 3814697265625ull;  // 5^18
        const int r = -(e + 18);
        unsigned __int128 n = 0;
        if (m != 0 && r < 127) {
            unsigned __int128 q = p >> r;
            const unsigned __int128 rem = p - (q << r);
            const unsigned __int128 half = (unsigned __int128)1 << (r - 1);
            if (rem > half || (rem == half && (q & 1))) ++q;
            n = q;
        }
        const uint64_t ip = uint64_t(n / 1000000000000000000ull);
        uint64_t fp = uint64_t(n % 1000000000000000000ull);
        char digits[20];
        for (int k = 17; k >= 0; --k) {
            digits[k] = char('0' + fp % 10);
            fp /= 10;
        }
        digits[18] = 0;
        len = snprintf(buf, sizeof buf, "%s%llu.%s", neg ? "-" : "", (unsigned long long)ip, digits);
    }
    if (len > kWidth) {
        out.append(kWidth, '*');
        return;
    }
    out.append(kWidth - len, ' ');
    out.append(buf, len);
}

// Fortran Iw: right justified in w columns, w '*' when it does not fit.
void put_iw(std::string &out, long long v, int w)
{
    char buf[32];
    const int len = snprintf(buf, sizeof buf, "%lld", v);
    if (len > w) {
        out.append(w, '*');
        return;
    }
    out.append(w - len, ' ');
    out.append(buf, len);
}

// One record of list-directed integer output as the reference's Fortran
// runtime (LLVM flang) writes it: each item preceded by one blank, and a
// new record begun (again with a blank) when an item would take the
// record past 79 characters.
void put_list_ints(std::string &out, const std::vector<long long> &items)
{
    constexpr int kRecord = 79;
    int col = 0;
    char buf[32];
    for (long long v : items) {
        const int len = snprintf(buf, sizeof buf, "%lld", v);
        if (col > 0 && col + 1 + len > kRecord) {
            out += '\n';
            col = 0;
        }
        out += ' ';
        out.append(buf, len);
        col += 1 + len;
    }
    out += '\n';
}

// Fortran (A6,I1) / (A,I1) of a name and a tracer number: the digit, or
// '*' when the number has more than one character.
std::string tracer_name(const char *base, int k)
{
    std::string s(base);
    if (k >= 0 && k <= 9)
        s += char('0' + k);
    else
        s += '*';
    return s;
}

// Formats rows [0, n) with fmt(i, out) on up to `threads` threads and
// writes the pieces in order.
template <class F>
bool write_rows(FILE *f, long long n, F fmt)
{
    if (n <= 0) return true;
    unsigned hw = std::thread::hardware_concurrency();
    const long long max_threads = std::max(1u, std::min(hw ? hw : 1u, 16u));
    const long long threads = std::max(1LL, std::min(max_threads, n / 8192));
    std::vector<std::string> parts(threads);
    auto work = [&](long long t) {
        const long long lo = n * t / threads, hi = n * (t + 1) / threads;
        std::string &s = parts[t];
        s.reserve(size_t(hi - lo) * 80);
        for (long long i = lo; i < hi; ++i) fmt(i, s);
    };
    if (threads == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (long long t = 1; t < threads; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto &th : pool) th.join();
    }
    for (auto &s : parts)
        if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) return false;
    return true;
}

int fail(const std::string &msg)
{
    bve::context_free_error() = "bve_vtk_write_frame: " + msg;
    return BVE_ERR_ARGUMENT;
}

}  // namespace

extern "C" int bve_vtk_write_frame(const char *filename, int n_particles, int ld_particles,
                                   const double *particles_x, const double *particles_relvort,
                                   const double *particles_absvort, int n_tracer,
                                   const double *particles_tracer, int panel_kind, int n_panels,
                                   int n_active, int ld_panels, const int *vertices,
                                   const int *has_children, const double *panels_tracer,
                                   const double *panels_relvort, const double *panels_absvort)
{
    if (!filename) return fail("null file name");
    if (n_particles < 0 || n_panels < 0 || n_active < 0 || n_tracer < 0)
        return fail("negative size");
    if (panel_kind != kTriPanel && panel_kind != kQuadPanel && panel_kind != kVoronoiPanel)
        return fail("panel_kind must be 3 (triangles), 4 (quadrilaterals) or 10 (Voronoi)");
    if (n_particles > 0 && !particles_x) return fail("null particle positions");
    if (n_panels > 0 && !vertices) return fail("null panel vertices");
    // nTracers comes from the particles (ParticlesEdgesPanels2.f90:2139-2141)
    const int n_tr = particles_tracer ? n_tracer : 0;
    if (n_tr > 0 && ld_particles < n_particles) return fail("ld_particles < n_particles");
    if (n_tr > 0 && panels_tracer && ld_panels < n_panels) return fail("ld_panels < n_panels");

    const long long np = n_particles, npan = n_panels;
    auto divided = [&](long long j) { return has_children && has_children[j] != 0; };
    std::vector<int> leaves;
    leaves.reserve(size_t(npan));
    for (long long j = 0; j < npan; ++j)
        if (!divided(j)) leaves.push_back(int(j));

    long long cell_list = 0;
    if (panel_kind != kVoronoiPanel) {
        cell_list = (long long)(panel_kind + 1) * n_active;
    } else {
        for (long long j = 0; j < npan; ++j)
            for (int k = 0; k < kVoronoiPanel; ++k)
                if (vertices[kVoronoiPanel * j + k] > 0) ++cell_list;
        cell_list += 2LL * n_active;
    }

    FILE *f = std::fopen(filename, "w");
    if (!f) {
        bve::context_free_error() = std::string("bve_vtk_write_frame: cannot open '") + filename +
                                    "': " + std::strerror(errno);
        return BVE_ERR_IO;
    }

    bool ok = true;
    std::string h;
    h += "# vtk DataFile Version 2.0\n";
    h += "Sphere Grid : ";
    put_iw(h, n_active, 8);
    h += " panels.\nASCII\nDATASET POLYDATA\nPOINTS ";
    put_iw(h, np, 8);
    h += "   double\n";
    ok = ok && std::fwrite(h.data(), 1, h.size(), f) == h.size();

    ok = ok && write_rows(f, np, [&](long long j, std::string &s) {
        for (int c = 0; c < 3; ++c) put_f24_18(s, particles_x[3 * j + c]);
        s += '\n';
    });

    h = "POLYGONS";
    put_iw(h, n_active, 8);
    put_iw(h, cell_list, 8);
    h += '\n';
    ok = ok && std::fwrite(h.data(), 1, h.size(), f) == h.size();
    if (panel_kind != kVoronoiPanel) {
        ok = ok && write_rows(f, (long long)leaves.size(), [&](long long i, std::string &s) {
            const long long j = leaves[i];
            put_iw(s, panel_kind, 8);
            for (int c = 0; c < panel_kind; ++c)
                put_iw(s, (long long)vertices[(long long)panel_kind * j + c] - 1, 8);
            s += '\n';
        });
    } else {
        // every panel, divided or not, as the reference's loop does
        ok = ok && write_rows(f, npan, [&](long long j, std::string &s) {
            int nv = 0;
            for (int k = 0; k < kVoronoiPanel; ++k)
                if (vertices[kVoronoiPanel * j + k] > 0) ++nv;
            std::vector<long long> items{nv};
            for (int k = 0; k < nv; ++k) items.push_back((long long)vertices[kVoronoiPanel * j + k] - 1);
            put_list_ints(s, items);
        });
    }

    auto scalars = [&](const std::string &header, long long n, auto value) {
        std::string s = header + "LOOKUP_TABLE default\n";
        ok = ok && std::fwrite(s.data(), 1, s.size(), f) == s.size();
        ok = ok && write_rows(f, n, [&](long long i, std::string &o) {
            put_f24_18(o, value(i));
            o += '\n';
        });
    };

    h = "POINT_DATA   ";
    put_iw(h, np, 8);
    h += '\n';
    ok = ok && std::fwrite(h.data(), 1, h.size(), f) == h.size();
    for (int k = 0; k < n_tr; ++k) {
        const double *col = particles_tracer + (long long)k * ld_particles;
        scalars(" SCALARS  " + tracer_name("Tracer", k + 1) + "  double 1\n", np,
                [&](long long j) { return col[j]; });
    }
    if (particles_relvort)
        scalars("SCALARS   relVort  double 1\n", np, [&](long long j) { return particles_relvort[j]; });
    if (particles_absvort)
        scalars("SCALARS   absVort  double 1\n", np, [&](long long j) { return particles_absvort[j]; });

    h = "CELL_DATA   ";
    put_iw(h, n_active, 8);
    h += '\n';
    ok = ok && std::fwrite(h.data(), 1, h.size(), f) == h.size();
    const long long nl = (long long)leaves.size();
    if (panels_tracer) {
        for (int k = 0; k < n_tr; ++k) {
            const double *col = panels_tracer + (long long)k * ld_panels;
            scalars("SCALARS " + tracer_name("TracerPanel", k + 1) + " double 1\n", nl,
                    [&](long long i) { return col[leaves[i]]; });
        }
    }
    if (panels_relvort)
        scalars("SCALARS  relVortPanel  double 1\n", nl,
                [&](long long i) { return panels_relvort[leaves[i]]; });
    if (panels_absvort)
        scalars("SCALARS  absVortPanel  double 1\n", nl,
                [&](long long i) { return panels_absvort[leaves[i]]; });

    if (std::fclose(f) != 0) ok = false;
    if (!ok) {
        bve::context_free_error() =
            std::string("bve_vtk_write_frame: write to '") + filename + "' failed: " + std::strerror(errno);
        return BVE_ERR_IO;
    }
    return BVE_OK;
}
