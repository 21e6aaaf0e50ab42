// Output files of a run: particle movements as JSON, the deposition histogram
// along a band and the per-triangle deposition (KDE input) file.  The formats
// are those of ParticleMovementTracker::saveMovementsToJson and
// SputteringModel::_writeHistogramToFile / _writeKDEToFile; the C interface is
// in include/start_output.h.
#include "start_output.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(const std::string& msg)
{
    g_error = msg;
    return -1;
}

struct P2 {
    double x, y;
};

// The centroid the reference prints and bins by (TriangleCell.cpp:11-14: the
// third vertex it passes is vertex(3), i.e. vertex(0)), or the true one.
void centroid(const double* t, bool reference, double& cx, double& cy)
{
    const double* v2 = reference ? t : t + 6;
    cx = (t[0] + t[3] + v2[0]) / 3.0;
    cy = (t[1] + t[4] + v2[1]) / 3.0;
}

double shoelace(const std::vector<P2>& p)
{
    double s = 0.0;
    for (size_t i = 0; i < p.size(); ++i) {
        const P2& a = p[i];
        const P2& b = p[(i + 1) % p.size()];
        s += a.x * b.y - b.x * a.y;
    }
    return std::fabs(s) * 0.5;
}

// Keeps the part of a convex polygon on one side of the vertical line x = x0
// (keep_right: x >= x0, else x <= x0); Sutherland-Hodgman for one edge.
std::vector<P2> clip_x(const std::vector<P2>& in, double x0, bool keep_right)
{
    std::vector<P2> out;
    const size_t n = in.size();
    for (size_t i = 0; i < n; ++i) {
        const P2& a = in[i];
        const P2& b = in[(i + 1) % n];
        const bool ka = keep_right ? a.x >= x0 : a.x <= x0;
        const bool kb = keep_right ? b.x >= x0 : b.x <= x0;
        if (ka) out.push_back(a);
        if (ka != kb) out.push_back(P2{x0, a.y + (b.y - a.y) * ((x0 - a.x) / (b.x - a.x))});
    }
    return out;
}

// nlohmann::json's number text (what dump() prints): the shortest digits
// d1..dk that read back to the same double, with the decimal point at n
// (value = 0.d1...dk * 10^n); plain notation for -4 < n <= 15 (".0" added to
// whole numbers), d.ddde+XX otherwise, at least two exponent digits.
void put_number(std::string& out, double v)
{
    if (!std::isfinite(v)) {  // JSON has no NaN or infinity; nlohmann writes null
        out += "null";
        return;
    }
    if (v == 0.0) {
        out += std::signbit(v) ? "-0.0" : "0.0";
        return;
    }
    if (v < 0.0) {
        out += '-';
        v = -v;
    }
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf - 1, v, std::chars_format::scientific);
    *r.ptr = '\0';  // to_chars does not terminate; atoi reads the exponent up to here
    const char* e = std::find(buf, r.ptr, 'e');
    std::string digits;
    for (const char* p = buf; p < e; ++p)
        if (*p != '.') digits += *p;
    const int k = (int)digits.size();
    const int n = std::atoi(e + 1) + 1;
    if (k <= n && n <= 15) {
        out += digits;
        out.append((size_t)(n - k), '0');
        out += ".0";
    } else if (0 < n && n <= 15) {
        out.append(digits, 0, (size_t)n);
        out += '.';
        out.append(digits, (size_t)n, std::string::npos);
    } else if (-4 < n && n <= 0) {
        out += "0.";
        out.append((size_t)(-n), '0');
        out += digits;
    } else {
        out += digits[0];
        if (k > 1) {
            out += '.';
            out.append(digits, 1, std::string::npos);
        }
        int x = n - 1;
        out += x < 0 ? "e-" : "e+";
        x = std::abs(x);
        if (x < 10) out += '0';
        out += std::to_string(x);
    }
}

}  // namespace

extern "C" {

const char* start_output_last_error(void) { return g_error.c_str(); }

int start_write_movements_json(const char* path, int64_t n, const uint64_t* ids, const int32_t* counts,
                               int32_t max_points, const double* xyz)
{
    if (!path || n < 0 || (n > 0 && (!ids || !counts || !xyz || max_points < 1)))
        return fail("invalid arguments");
    if (n == 0) return 0;  // an empty map: nothing written (ParticleMovementTracker.cpp:24-28)
    // the reference's json object: keys in string order, the last of equal ids kept
    std::map<std::string, int64_t> order;
    for (int64_t s = 0; s < n; ++s)
        if (counts[s] > 1) order[std::to_string(ids[s])] = s;
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(std::string("cannot open ") + path + " for writing");
    if (order.empty()) {
        f << "null";
        return f ? 0 : fail("write failed");
    }
    std::string out = "{\n";
    bool first = true;
    for (const auto& kv : order) {
        const int64_t s = kv.second;
        const int32_t m = counts[s] < max_points ? counts[s] : max_points;
        if (!first) out += ",\n";
        first = false;
        out += "    \"" + kv.first + "\": [\n";
        for (int32_t k = 0; k < m; ++k) {
            const double* q = xyz + 3 * ((int64_t)s * max_points + k);
            out += "        {\n            \"x\": ";
            put_number(out, q[0]);
            out += ",\n            \"y\": ";
            put_number(out, q[1]);
            out += ",\n            \"z\": ";
            put_number(out, q[2]);
            out += k + 1 < m ? "\n        },\n" : "\n        }\n";
        }
        out += "    ]";
        if (out.size() > (1u << 20)) {
            f << out;
            out.clear();
        }
    }
    out += "\n}";
    f << out;
    f.close();
    return f ? 0 : fail(std::string("write failed: ") + path);
}

double start_band_overlap(const double* t, double x_min, double x_max)
{
    std::vector<P2> tri{{t[0], t[1]}, {t[3], t[4]}, {t[6], t[7]}};
    const double area = shoelace(tri);
    if (!(area > 0.0)) return 0.0;
    std::vector<P2> p = clip_x(tri, x_min, true);
    if (p.size() >= 3) p = clip_x(p, x_max, false);
    return p.size() >= 3 ? shoelace(p) / area : 0.0;
}

int start_deposition_histogram(int64_t n_tris, const double* tri_xyz, const uint64_t* counts, double x_min,
                               double x_max, double bin, int32_t reference_centroid, int64_t* n_bins,
                               double* bin_y, double* bin_total)
{
    if (n_tris < 0 || (n_tris > 0 && (!tri_xyz || !counts)) || !n_bins || !(bin > 0.0))
        return fail("invalid arguments");
    std::map<double, double> bins;
    for (int64_t i = 0; i < n_tris; ++i) {
        const double* t = tri_xyz + 9 * i;
        double cx, cy;
        centroid(t, reference_centroid != 0, cx, cy);
        if (cx < x_min || cx > x_max) continue;
        const double overlap = start_band_overlap(t, x_min, x_max);
        if (overlap >= 0.75) bins[std::floor(cy / bin) * bin] += (double)counts[i] * overlap;
    }
    const int64_t cap = *n_bins;
    *n_bins = (int64_t)bins.size();
    if (cap == 0) return 0;
    if (cap < (int64_t)bins.size() || !bin_y || !bin_total) return fail("bin capacity too small");
    int64_t k = 0;
    for (const auto& kv : bins) {
        bin_y[k] = kv.first;
        bin_total[k] = kv.second;
        ++k;
    }
    return 0;
}

int start_write_histogram(const char* path, int64_t n_tris, const double* tri_xyz, const uint64_t* counts,
                          double x_min, double x_max, double bin, int32_t reference_centroid)
{
    if (!path) return fail("invalid arguments");
    int64_t nb = 0;
    if (start_deposition_histogram(n_tris, tri_xyz, counts, x_min, x_max, bin, reference_centroid, &nb, nullptr,
                                   nullptr) != 0)
        return -1;
    std::vector<double> y((size_t)nb + 1), tot((size_t)nb + 1);
    if (nb > 0 && start_deposition_histogram(n_tris, tri_xyz, counts, x_min, x_max, bin, reference_centroid, &nb,
                                             y.data(), tot.data()) != 0)
        return -1;
    std::ofstream f(path);
    if (!f) return fail(std::string("cannot open ") + path + " for writing");
    for (int64_t k = 0; k < nb; ++k) f << y[k] << " " << tot[k] << "\n";
    f.close();
    return f ? 0 : fail(std::string("write failed: ") + path);
}

int start_write_kde(const char* path, int64_t n_tris, const double* tri_xyz, const uint64_t* counts,
                    int32_t reference_centroid)
{
    if (!path || n_tris < 0 || (n_tris > 0 && (!tri_xyz || !counts))) return fail("invalid arguments");
    std::ofstream f(path);
    if (!f) return fail(std::string("cannot open ") + path + " for writing");
    f << "# X-coordinate(cm) Y-coordinate(cm) Count\n";
    for (int64_t i = 0; i < n_tris; ++i) {
        if (counts[i] == 0) continue;
        double cx, cy;
        centroid(tri_xyz + 9 * i, reference_centroid != 0, cx, cy);
        f << cx << " " << cy << " " << counts[i] << "\n";
    }
    f.close();
    return f ? 0 : fail(std::string("write failed: ") + path);
}

}  // extern "C"
