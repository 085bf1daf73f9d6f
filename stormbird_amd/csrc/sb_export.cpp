// Wake export in the reference's formats (wake/dynamic_wake/export.rs:16-210):
// Wavefront OBJ (points, one quad face per panel, 1-based), VTK XML PolyData
// (points, quad connectivity, offsets, per-panel strength and viscous core
// length) and the per-panel point indices a plotly mesh is built from.
//
// Numbers are written the way Rust's `{}` prints an f64: the shortest digits
// that read back to the same value (std::to_chars in scientific form without a
// precision gives them), placed in positional notation with zeros as needed
// (no exponent), "-0" for negative zero, "NaN" and "inf"/"-inf".
#include <charconv>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "sb_host.h"

namespace sb_host {

std::string rust_f64(double v) {
    if (v != v) return "NaN";
    if (v == __builtin_inf()) return "inf";
    if (v == -__builtin_inf()) return "-inf";
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    if (r.ec != std::errc()) throw std::runtime_error("number formatting failed");
    std::string t(buf, r.ptr);  // [-]d[.ddd]e(+|-)xx
    bool neg = t[0] == '-';
    if (neg) t.erase(0, 1);
    size_t e = t.find('e');
    int ex = std::stoi(t.substr(e + 1));
    std::string dg;
    for (size_t q = 0; q < e; ++q)
        if (t[q] != '.') dg += t[q];
    const int n = (int)dg.size();  // value = d.ddd x 10^ex
    std::string out;
    if (ex >= n - 1) out = dg + std::string(ex - (n - 1), '0');
    else if (ex >= 0) out = dg.substr(0, ex + 1) + "." + dg.substr(ex + 1);
    else out = "0." + std::string(-ex - 1, '0') + dg;
    return neg ? "-" + out : out;
}

// dynamic_wake/mod.rs:82-91: (s, i + wing), (s, i + 1 + wing), (s + 1, i + 1 + wing), (s + 1, i + wing)
void panel_point_indices(const sb_sim& s, long k, long out[4]) {
    long row = k / s.N;
    int i = (int)(k - row * s.N);
    long a = row * s.P + i + s.setup.lfm.wing_of[i];
    out[0] = a;
    out[1] = a + 1;
    out[2] = a + s.P + 1;
    out[3] = a + s.P;
}

namespace {

void download_wake(const sb_sim& s, std::vector<double>& pts, std::vector<double>& str) {
    if (s.setup.quasi) throw std::runtime_error("a quasi-steady simulation has no dynamic wake to export");
    check(hipSetDevice(s.device), "hipSetDevice");
    pts.resize(3 * s.nr_points);
    str.resize(s.nr_panels);
    check(hipStreamSynchronize(s.stream), "sync");
    check(hipMemcpy(pts.data(), s.d_pts[s.cur], sizeof(double) * pts.size(), hipMemcpyDeviceToHost), "download");
    check(hipMemcpy(str.data(), s.d_str[s.scur], sizeof(double) * str.size(), hipMemcpyDeviceToHost), "download");
}

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    f << text;
    if (!f) throw std::runtime_error("cannot write " + path);
}

}  // namespace

// export.rs:30-64
void write_wake_obj(const sb_sim& s, const std::string& path) {
    std::vector<double> pts, str;
    download_wake(s, pts, str);
    std::ostringstream o;
    o << "o wake\n";
    for (long q = 0; q < s.nr_points; ++q)
        o << "v " << rust_f64(pts[3 * q]) << " " << rust_f64(pts[3 * q + 1]) << " " << rust_f64(pts[3 * q + 2]) << "\n";
    long ix[4];
    for (long k = 0; k < s.nr_panels; ++k) {
        panel_point_indices(s, k, ix);
        o << "f " << ix[0] + 1 << " " << ix[1] + 1 << " " << ix[2] + 1 << " " << ix[3] + 1 << "\n";
    }
    write_file(path, o.str());
}

// export.rs:66-159
void write_wake_vtk(const sb_sim& s, const std::string& path) {
    std::vector<double> pts, str;
    download_wake(s, pts, str);
    const std::vector<double>& core = s.setup.wake.core;
    std::ostringstream o;
    o << "<?xml version=\"1.0\"?>\n";
    o << "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    o << "\t<PolyData>\n";
    o << "\t\t<Piece NumberOfPoints=\"" << s.nr_points
      << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"" << s.nr_panels << "\">\n";
    o << "\t\t\t<Points>\n";
    o << "\t\t\t\t<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (long q = 0; q < s.nr_points; ++q)
        o << "\t\t\t\t\t" << rust_f64(pts[3 * q]) << " " << rust_f64(pts[3 * q + 1]) << " " << rust_f64(pts[3 * q + 2])
          << "\n";
    o << "\t\t\t\t</DataArray>\n";
    o << "\t\t\t</Points>\n";
    o << "\t\t\t<Polys>\n";
    o << "\t\t\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    long ix[4];
    for (long k = 0; k < s.nr_panels; ++k) {
        panel_point_indices(s, k, ix);
        o << "\t\t\t\t\t" << ix[0] << " " << ix[1] << " " << ix[2] << " " << ix[3] << "\n";
    }
    o << "\t\t\t\t</DataArray>\n";
    o << "\t\t\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (long k = 0; k < s.nr_panels; ++k) o << "\t\t\t\t\t" << (k + 1) * 4 << "\n";
    o << "\t\t\t\t</DataArray>\n";
    o << "\t\t\t</Polys>\n";
    o << "\t\t\t<CellData Scalars=\"strength\">\n";
    o << "\t\t\t\t<DataArray type=\"Float32\" Name=\"strength\" format=\"ascii\">\n";
    for (long k = 0; k < s.nr_panels; ++k) o << "\t\t\t\t\t" << rust_f64(str[k]) << "\n";
    o << "\t\t\t\t</DataArray>\n";
    o << "\t\t\t\t<DataArray type=\"Float32\" Name=\"viscous_core_length\" format=\"ascii\">\n";
    for (long k = 0; k < s.nr_panels; ++k) o << "\t\t\t\t\t" << rust_f64(core[k]) << "\n";
    o << "\t\t\t\t</DataArray>\n";
    o << "\t\t\t</CellData>\n";
    o << "\t\t</Piece>\n";
    o << "\t</PolyData>\n";
    o << "</VTKFile>\n";
    write_file(path, o.str());
}

// export.rs:16-28 with simulation.rs:183-185: `{folder}/wake_{(time / time_step) as usize}.vtp`
// (the cast saturates: NaN and negative values give 0).
void write_wake_step_file(const sb_sim& s, double time, double dt) {
    double q = time / dt;
    unsigned long long idx = 0;
    if (q == q && q > 0.0) idx = q >= 18446744073709551615.0 ? ~0ull : (unsigned long long)q;
    write_wake_vtk(s, s.setup.wake.files_folder + "/wake_" + std::to_string(idx) + ".vtp");
}

}  // namespace sb_host

extern "C" {

int sb_sim_write_wake_obj(const sb_sim* s, const char* path) {
    if (!s || !path) return -1;
    return guarded([&] { write_wake_obj(*s, path); });
}

int sb_sim_write_wake_vtk(const sb_sim* s, const char* path) {
    if (!s || !path) return -1;
    return guarded([&] { write_wake_vtk(*s, path); });
}

int sb_sim_get_panel_point_indices(const sb_sim* s, long* out, size_t cap) {
    if (!s || !out) return -1;
    if (cap < 4 * (size_t)s->nr_panels) return -3;
    for (long k = 0; k < s->nr_panels; ++k) panel_point_indices(*s, k, out + 4 * k);
    return 0;
}

int sb_debug_format_f64(double v, char* buf, size_t cap) {
    if (!buf || cap == 0) return -1;
    std::string t = rust_f64(v);
    if (t.size() + 1 > cap) return -3;
    memcpy(buf, t.c_str(), t.size() + 1);
    return (int)t.size();
}

}  // extern "C"
