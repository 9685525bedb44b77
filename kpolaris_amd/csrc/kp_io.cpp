// kp_io.cpp -- HDF5 input and output around the imaging ABI (include/kp_io.h):
// iHARM dumps as the reference's loader reads them, the kpolaris_image v1
// product as the reference's image tool writes it, and the publication norms
// of the reference's Stokes comparison script.  Host code over the HDF5 C API;
// no device code.
#include "../../include/kp_io.h"
#include "kp_errors.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

using kp_internal::fail;

// An hid_t closed with its own H5*close on scope exit.
struct Handle {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Handle() = default;
    Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    Handle(Handle&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    ~Handle() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

// HDF5 prints its error stack by default; the status and kp_last_error()
// carry the failure instead.
struct QuietErrors {
    H5E_auto2_t func = nullptr;
    void* data = nullptr;
    QuietErrors() {
        H5Eget_auto2(H5E_DEFAULT, &func, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietErrors() { H5Eset_auto2(H5E_DEFAULT, func, data); }
};

bool exists(hid_t loc, const std::string& path) {
    // each intermediate link must exist before the next is queried
    size_t pos = 0;
    while (true) {
        pos = path.find('/', pos + 1);
        const std::string part = path.substr(0, pos);
        if (!part.empty() && part != "/" && H5Lexists(loc, part.c_str(), H5P_DEFAULT) <= 0) return false;
        if (pos == std::string::npos) return true;
    }
}

bool read_scalar(hid_t file, const std::string& path, double& out) {
    if (!exists(file, path)) return false;
    Handle d(H5Dopen2(file, path.c_str(), H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return false;
    return H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &out) >= 0;
}

bool read_scalar_int(hid_t file, const std::string& path, int32_t& out) {
    if (!exists(file, path)) return false;
    Handle d(H5Dopen2(file, path.c_str(), H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return false;
    return H5Dread(d, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, &out) >= 0;
}

bool write_scalar(hid_t loc, const char* name, double v) {
    Handle s(H5Screate(H5S_SCALAR), H5Sclose);
    Handle d(H5Dcreate2(loc, name, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) >= 0;
}

bool write_scalar_int(hid_t loc, const char* name, int32_t v) {
    Handle s(H5Screate(H5S_SCALAR), H5Sclose);
    Handle d(H5Dcreate2(loc, name, H5T_STD_I32LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) >= 0;
}

bool attr_string(hid_t loc, const char* name, const std::string& v) {
    Handle t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(t, std::max<size_t>(v.size(), 1));
    H5Tset_strpad(t, H5T_STR_NULLPAD);
    Handle s(H5Screate(H5S_SCALAR), H5Sclose);
    Handle a(H5Acreate2(loc, name, t, s, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    std::string buf = v.empty() ? std::string(1, '\0') : v;
    return a.ok() && H5Awrite(a, t, buf.data()) >= 0;
}

bool attr_double(hid_t loc, const char* name, double v) {
    Handle s(H5Screate(H5S_SCALAR), H5Sclose);
    Handle a(H5Acreate2(loc, name, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a, H5T_NATIVE_DOUBLE, &v) >= 0;
}

bool attr_int(hid_t loc, const char* name, int32_t v) {
    Handle s(H5Screate(H5S_SCALAR), H5Sclose);
    Handle a(H5Acreate2(loc, name, H5T_STD_I32LE, s, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a, H5T_NATIVE_INT32, &v) >= 0;
}

// String attribute of either fixed or variable length; false if absent.
bool read_attr_string(hid_t loc, const char* name, std::string& out) {
    if (H5Aexists(loc, name) <= 0) return false;
    Handle a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    Handle t(H5Aget_type(a), H5Tclose);
    if (H5Tget_class(t) != H5T_STRING) return false;
    if (H5Tis_variable_str(t) > 0) {
        Handle mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        char* p = nullptr;
        if (H5Aread(a, mt, &p) < 0 || !p) return false;
        out = p;
        H5free_memory(p);
        return true;
    }
    const size_t n = H5Tget_size(t);
    std::vector<char> buf(n + 1, '\0');
    if (H5Aread(a, t, buf.data()) < 0) return false;
    out.assign(buf.data(), strnlen(buf.data(), n));
    return true;
}

bool read_attr_double(hid_t loc, const char* name, double& out) {
    if (H5Aexists(loc, name) <= 0) return false;
    Handle a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    return H5Aread(a, H5T_NATIVE_DOUBLE, &out) >= 0;
}

bool write_1d(hid_t loc, const char* name, const double* v, size_t n) {
    hsize_t dims[1] = {(hsize_t)n};
    Handle s(H5Screate_simple(1, dims, nullptr), H5Sclose);
    Handle d(H5Dcreate2(loc, name, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && (n == 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0);
}

// Image arrays: HDF5 shape (ny, nx), row-major [iy, ix].
bool write_2d(hid_t loc, const char* name, const double* v, int nx, int ny) {
    hsize_t dims[2] = {(hsize_t)ny, (hsize_t)nx};
    Handle s(H5Screate_simple(2, dims, nullptr), H5Sclose);
    Handle d(H5Dcreate2(loc, name, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
}

bool write_2d_int(hid_t loc, const char* name, const int32_t* v, int nx, int ny) {
    hsize_t dims[2] = {(hsize_t)ny, (hsize_t)nx};
    Handle s(H5Screate_simple(2, dims, nullptr), H5Sclose);
    Handle d(H5Dcreate2(loc, name, H5T_STD_I32LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
}

// A dataset of any rank read whole as doubles (or int64 for the integer
// diagnostics); shape returned in dims.
template <class T>
bool read_array(hid_t loc, const std::string& path, hid_t memtype, std::vector<T>& out, std::vector<hsize_t>& dims) {
    if (!exists(loc, path)) return false;
    Handle d(H5Dopen2(loc, path.c_str(), H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return false;
    Handle s(H5Dget_space(d), H5Sclose);
    const int rank = H5Sget_simple_extent_ndims(s);
    if (rank < 0) return false;
    dims.assign((size_t)rank, 0);
    H5Sget_simple_extent_dims(s, dims.data(), nullptr);
    size_t n = 1;
    for (hsize_t x : dims) n *= (size_t)x;
    out.assign(n, T(0));
    return n == 0 || H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
}

std::string fmt17(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

}  // namespace

// ===========================================================================
// iHARM dumps (src/grmhd/iharm_loader.cpp:262-396)
// ===========================================================================
extern "C" kp_status kp_iharm_read_header(const char* path, kp_iharm_header* h) {
    kp_internal::clear_error();
    if (!path || !h) return fail(KP_INVALID_ARGUMENT, "null argument");
    QuietErrors quiet;
    Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(KP_INVALID_ARGUMENT, std::string("cannot open iharm dump ") + path);
    std::memset(h, 0, sizeof(*h));
    const auto need = [&](const std::string& p, double& v) {
        return read_scalar(f, p, v) ? KP_OK : fail(KP_INVALID_ARGUMENT, "iharm dump lacks " + p);
    };
    if (!read_scalar_int(f, "/header/n1", h->n1) || !read_scalar_int(f, "/header/n2", h->n2) ||
        !read_scalar_int(f, "/header/n3", h->n3))
        return fail(KP_INVALID_ARGUMENT, "iharm dump lacks /header/n1..n3");
    kp_status st;
    if ((st = need("/header/a", h->a)) != KP_OK) return st;
    if (!read_scalar(f, "/header/gam", h->gam) && !read_scalar(f, "/header/gamma", h->gam))
        return fail(KP_INVALID_ARGUMENT, "iharm dump lacks /header/gam and /header/gamma");
    const char* geom[6] = {"startx1", "startx2", "startx3", "dx1", "dx2", "dx3"};
    double* gv[6] = {&h->startx1, &h->startx2, &h->startx3, &h->dx1, &h->dx2, &h->dx3};
    for (int i = 0; i < 6; ++i)
        if ((st = need(std::string("/header/geom/") + geom[i], *gv[i])) != KP_OK) return st;
    const auto finite = [](double v) { return std::isfinite(v); };
    if (h->n1 <= 0 || h->n2 <= 0 || h->n3 <= 0) return fail(KP_INVALID_ARGUMENT, "iharm header dimensions must be positive");
    if (!finite(h->a) || std::abs(h->a) > 1.0)
        return fail(KP_INVALID_ARGUMENT, "iharm header spin must be finite and lie in [-1, 1]");
    if (!finite(h->gam) || h->gam <= 1.0)
        return fail(KP_INVALID_ARGUMENT, "iharm header gamma must be finite and greater than 1");
    if (!finite(h->startx1) || !finite(h->startx2) || !finite(h->startx3))
        return fail(KP_INVALID_ARGUMENT, "iharm grid starts must be finite");
    if (!finite(h->dx1) || !finite(h->dx2) || !finite(h->dx3) || h->dx1 <= 0.0 || h->dx2 <= 0.0 || h->dx3 <= 0.0)
        return fail(KP_INVALID_ARGUMENT, "iharm grid spacings must be finite and positive");
    std::string g = "/header/geom/fmks";
    if (!exists(f, g + "/hslope")) g = "/header/geom/mmks";
    const char* mk[6] = {"hslope", "mks_smooth", "poly_alpha", "poly_xt", "r_in", "r_out"};
    double* mv[6] = {&h->hslope, &h->mks_smooth, &h->poly_alpha, &h->poly_xt, &h->r_in, &h->r_out};
    for (int i = 0; i < 6; ++i)
        if ((st = need(g + "/" + mk[i], *mv[i])) != KP_OK) return st;
    if (!finite(h->hslope) || !finite(h->mks_smooth) || !finite(h->poly_alpha) || !finite(h->poly_xt) ||
        h->poly_alpha <= -1.0 || h->poly_xt <= 0.0)
        return fail(KP_INVALID_ARGUMENT,
                    "iharm FMKS/MMKS mapping parameters must be finite with poly_alpha > -1 and poly_xt > 0");
    if (!finite(h->r_in) || !finite(h->r_out) || h->r_in <= 0.0 || h->r_out <= h->r_in)
        return fail(KP_INVALID_ARGUMENT, "iharm radial bounds must be finite, positive, and increasing");
    const double pi = 3.141592653589793238462643383279502884;
    h->poly_norm = 0.5 * pi / (1.0 + 1.0 / (h->poly_alpha + 1.0) / std::pow(h->poly_xt, h->poly_alpha));
    if (!finite(h->poly_norm) || h->poly_norm <= 0.0)
        return fail(KP_INVALID_ARGUMENT, "iharm FMKS/MMKS mapping normalization must be finite and positive");
    // the primitive count from /prims
    if (!exists(f, "/prims")) return fail(KP_INVALID_ARGUMENT, "iharm dump lacks /prims");
    Handle d(H5Dopen2(f, "/prims", H5P_DEFAULT), H5Dclose);
    Handle s(H5Dget_space(d), H5Sclose);
    if (H5Sget_simple_extent_ndims(s) != 4) return fail(KP_INVALID_ARGUMENT, "iharm /prims dataset must be rank 4");
    hsize_t dims[4] = {0, 0, 0, 0};
    H5Sget_simple_extent_dims(s, dims, nullptr);
    if ((int64_t)dims[0] != h->n1 || (int64_t)dims[1] != h->n2 || (int64_t)dims[2] != h->n3 || dims[3] < 8)
        return fail(KP_INVALID_ARGUMENT,
                    "iharm /prims dimensions do not match header or have fewer than 8 primitives");
    const uint64_t cells = (uint64_t)dims[0] * dims[1] * dims[2];
    if (cells > (uint64_t)std::numeric_limits<int>::max())
        return fail(KP_INVALID_ARGUMENT, "iharm grid cell count exceeds the supported int index range");
    h->nprim = (int32_t)dims[3];
    return KP_OK;
}

extern "C" kp_status kp_iharm_read_cells(const char* path, const kp_iharm_header* h, kp_grmhd_cell* cells) {
    kp_internal::clear_error();
    if (!path || !h || !cells) return fail(KP_INVALID_ARGUMENT, "null argument");
    QuietErrors quiet;
    Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(KP_INVALID_ARGUMENT, std::string("cannot open iharm dump ") + path);
    std::vector<float> raw;
    std::vector<hsize_t> dims;
    if (!read_array(f, "/prims", H5T_NATIVE_FLOAT, raw, dims) || dims.size() != 4)
        return fail(KP_INVALID_ARGUMENT, "iharm /prims dataset must be rank 4");
    if ((int64_t)dims[0] != h->n1 || (int64_t)dims[1] != h->n2 || (int64_t)dims[2] != h->n3 ||
        (int64_t)dims[3] != h->nprim || dims[3] < 8)
        return fail(KP_INVALID_ARGUMENT, "iharm /prims dimensions do not match the header");
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    const size_t np = (size_t)dims[3];
    for (size_t c = 0; c < ncell; ++c) {
        for (int v = 0; v < 8; ++v) {
            const float x = raw[c * np + v];
            if (!std::isfinite((double)x))
                return fail(KP_INVALID_ARGUMENT, "iharm /prims first eight primitive fields must be finite");
            cells[c].prim[v] = x;
        }
        for (int v = 0; v < 5; ++v) cells[c].derived[v] = 0.0;
    }
    return KP_OK;
}

extern "C" kp_status kp_iharm_load_snapshot(void* user, int32_t index, kp_grmhd_cell* cells, uint64_t ncells) {
    kp_internal::clear_error();
    const kp_iharm_sequence* q = static_cast<const kp_iharm_sequence*>(user);
    if (!q || !q->paths || !cells) return fail(KP_INVALID_ARGUMENT, "null argument");
    if (index < 0 || index >= q->nsnap || !q->paths[index]) return fail(KP_INVALID_ARGUMENT, "snapshot index out of range");
    const kp_iharm_header& g = q->header;
    if (ncells != (uint64_t)g.n1 * (uint64_t)g.n2 * (uint64_t)g.n3)
        return fail(KP_INVALID_ARGUMENT, "snapshot cell count does not match the sequence grid");
    kp_iharm_header h;
    kp_status st = kp_iharm_read_header(q->paths[index], &h);
    if (st != KP_OK) return st;
    const bool same = h.n1 == g.n1 && h.n2 == g.n2 && h.n3 == g.n3 && h.a == g.a && h.startx1 == g.startx1 &&
                      h.startx2 == g.startx2 && h.startx3 == g.startx3 && h.dx1 == g.dx1 && h.dx2 == g.dx2 &&
                      h.dx3 == g.dx3 && h.hslope == g.hslope && h.mks_smooth == g.mks_smooth &&
                      h.poly_alpha == g.poly_alpha && h.poly_xt == g.poly_xt;
    if (!same)
        return fail(KP_INVALID_ARGUMENT,
                    "slow_light_interpolation=fluid requires native fluid variables, a common coordinate "
                    "mapping/domain, and matching scalar reconstruction");
    return kp_iharm_read_cells(q->paths[index], &h, cells);
}

extern "C" kp_status kp_iharm_write(const char* path, const kp_iharm_header* h, const float* prims) {
    kp_internal::clear_error();
    if (!path || !h || !prims) return fail(KP_INVALID_ARGUMENT, "null argument");
    if (h->n1 <= 0 || h->n2 <= 0 || h->n3 <= 0 || h->nprim < 8)
        return fail(KP_INVALID_ARGUMENT, "iharm dump needs positive dimensions and at least 8 primitives");
    QuietErrors quiet;
    Handle f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(KP_INVALID_ARGUMENT, std::string("cannot create ") + path);
    Handle hd(H5Gcreate2(f, "/header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    Handle geom(H5Gcreate2(hd, "geom", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    Handle fm(H5Gcreate2(geom, "fmks", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    bool ok = hd.ok() && geom.ok() && fm.ok();
    ok = ok && write_scalar_int(hd, "n1", h->n1) && write_scalar_int(hd, "n2", h->n2) &&
         write_scalar_int(hd, "n3", h->n3) && write_scalar(hd, "a", h->a) && write_scalar(hd, "gam", h->gam);
    ok = ok && write_scalar(geom, "startx1", h->startx1) && write_scalar(geom, "startx2", h->startx2) &&
         write_scalar(geom, "startx3", h->startx3) && write_scalar(geom, "dx1", h->dx1) &&
         write_scalar(geom, "dx2", h->dx2) && write_scalar(geom, "dx3", h->dx3);
    ok = ok && write_scalar(fm, "hslope", h->hslope) && write_scalar(fm, "mks_smooth", h->mks_smooth) &&
         write_scalar(fm, "poly_alpha", h->poly_alpha) && write_scalar(fm, "poly_xt", h->poly_xt) &&
         write_scalar(fm, "r_in", h->r_in) && write_scalar(fm, "r_out", h->r_out);
    hsize_t dims[4] = {(hsize_t)h->n1, (hsize_t)h->n2, (hsize_t)h->n3, (hsize_t)h->nprim};
    Handle s(H5Screate_simple(4, dims, nullptr), H5Sclose);
    Handle d(H5Dcreate2(f, "/prims", H5T_IEEE_F32LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    ok = ok && d.ok() && H5Dwrite(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, prims) >= 0;
    if (!ok) return fail(KP_HIP_ERROR, std::string("HDF5 write of ") + path + " failed");
    return KP_OK;
}

extern "C" void kp_iharm_apply_header(const kp_iharm_header* h, kp_grmhd_model* m) {
    if (!h || !m) return;
    m->n1 = h->n1;
    m->n2 = h->n2;
    m->n3 = h->n3;
    m->spin = h->a;
    m->gam = h->gam;
    m->startx1 = h->startx1;
    m->startx2 = h->startx2;
    m->startx3 = h->startx3;
    m->dx1 = h->dx1;
    m->dx2 = h->dx2;
    m->dx3 = h->dx3;
    m->hslope = h->hslope;
    m->mks_smooth = h->mks_smooth;
    m->poly_alpha = h->poly_alpha;
    m->poly_xt = h->poly_xt;
    m->poly_norm = h->poly_norm;
    m->r_in = h->r_in;
    m->r_out = h->r_out;
}

// ===========================================================================
// kpolaris_image v1 (docs/hdf5_schema.md:81-200;
// tools/kpolaris_model_image.cpp:4283-4880)
// ===========================================================================
namespace {

bool common_attrs(hid_t loc, const kp_image_file_meta& m, const kp_pass_a_params& p, const std::string& flist) {
    bool ok = attr_string(loc, "schema", "kpolaris_image") && attr_int(loc, "schema_version", 1) &&
              attr_string(loc, "code", "KPolaris") && attr_string(loc, "code_version", "kpolaris_amd") &&
              attr_string(loc, "code_revision", "unknown") && attr_int(loc, "code_source_dirty", -1) &&
              attr_string(loc, "code_source_fingerprint", "unknown") &&
              attr_string(loc, "compiler_id", "hipcc") && attr_string(loc, "compiler_version", "rocm") &&
              attr_string(loc, "build_type", "Release");
    ok = ok && attr_string(loc, "model", m.model ? m.model : "riaf") &&
         attr_string(loc, "camera", m.camera ? m.camera : "pinhole") &&
         attr_string(loc, "coordinate", m.coordinate ? m.coordinate : "bl") &&
         attr_string(loc, "stokes_convention", "camera_frame") &&
         attr_string(loc, "polarization_basis", "camera_screen") && attr_string(loc, "evpa_0", m.evpa_0) &&
         attr_int(loc, "nx", p.camera.nx) && attr_int(loc, "ny", p.camera.ny) && attr_int(loc, "nfreq", m.nfreq) &&
         attr_string(loc, "frequency_list_hz", flist) && attr_double(loc, "frequency_hz", m.frequency_hz[0]);
    if (m.parameter_file) ok = ok && attr_string(loc, "parameter_file", m.parameter_file);
    return ok;
}

}  // namespace

extern "C" kp_status kp_image_write_hdf5(const char* path, const kp_image_file_meta* meta,
                                         const kp_pass_a_params* params, const kp_image_data* img) {
    kp_internal::clear_error();
    if (!path || !meta || !params || !img) return fail(KP_INVALID_ARGUMENT, "null argument");
    const kp_image_file_meta& m = *meta;
    const kp_pass_a_params& p = *params;
    const int nx = p.camera.nx, ny = p.camera.ny;
    if (nx <= 0 || ny <= 0) return fail(KP_INVALID_ARGUMENT, "camera nx and ny must be positive");
    if (m.nfreq < 1 || !m.frequency_hz) return fail(KP_INVALID_ARGUMENT, "at least one frequency required");
    const std::string evpa = m.evpa_0 ? m.evpa_0 : "N";
    if (evpa != "N" && evpa != "W") return fail(KP_INVALID_ARGUMENT, "evpa_0 must be N or W");
    const double* real[11] = {img->closure_x, img->closure_k, img->final_null, img->frame_error, img->det_r,
                              img->overlap_r11, img->overlap_r12, img->overlap_r21, img->overlap_r22,
                              img->basis_identity_error, img->basis_rotation_angle};
    if (!img->image_i || !img->image_q || !img->image_u || !img->image_v || !img->reason || !img->steps ||
        !img->pass_a_steps)
        return fail(KP_INVALID_ARGUMENT, "Stokes arrays, reason and step counts are required");
    for (const double* r : real)
        if (!r) return fail(KP_INVALID_ARGUMENT, "every per-pixel diagnostic array is required");
    const size_t np = (size_t)nx * (size_t)ny;
    const int qu_sign = evpa == "N" ? -1 : 1;
    const double scale = m.intensity_to_flux_jy_per_pixel;
    const bool scale_valid = std::isfinite(scale) && scale > 0.0;
    std::string flist;
    for (int f = 0; f < m.nfreq; ++f) flist += (f ? "," : "") + fmt17(m.frequency_hz[f]);

    QuietErrors quiet;
    Handle file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(KP_INVALID_ARGUMENT, std::string("cannot create ") + path);
    // camera extents as the kernels use them (x_extent = 2 xspan, or fov)
    const kp_camera_params& c = p.camera;
    const double xext = c.xspan > 0.0 ? 2.0 * c.xspan : c.fov;
    const double yext = c.yspan > 0.0 ? 2.0 * c.yspan
                        : (c.xspan > 0.0 ? 2.0 * c.xspan * double(ny) / double(nx) : (c.fov_y > 0.0 ? c.fov_y : c.fov));
    const double xh = 0.5 * xext, yh = 0.5 * yext;
    const double dx = xext / nx, dy = yext / ny;
    bool ok = common_attrs(file, m, p, flist);
    ok = ok && attr_string(file, "metric", m.metric ? m.metric : "kerr") &&
         attr_double(file, "camera_radius", c.radius) && attr_double(file, "inclination_rad", c.inclination) &&
         attr_double(file, "fov", c.fov) && attr_double(file, "fovy", c.fov_y > 0.0 ? c.fov_y : c.fov) &&
         attr_double(file, "dsource", m.dsource_pc) && attr_double(file, "mbh_solar", m.mbh_solar) &&
         attr_double(file, "x_offset", c.x_offset) && attr_double(file, "y_offset", c.y_offset) &&
         attr_string(file, "grid_units", "dimensionless_slope") && attr_double(file, "x_half_extent", xh) &&
         attr_double(file, "y_half_extent", yh) && attr_double(file, "dx", dx) && attr_double(file, "dy", dy) &&
         attr_int(file, "flux_scale_valid", scale_valid ? 1 : 0) &&
         attr_double(file, "intensity_to_flux_jy_per_pixel", scale) &&
         attr_string(file, "image_stokes_units", "cgs_specific_intensity") &&
         attr_string(file, "flux_density_unit", "Jy") && attr_double(file, "spin", m.spin) &&
         attr_string(file, "image_array_order", "y,x") && attr_string(file, "image_display_origin", "lower");
    {
        Handle h(H5Gcreate2(file, "/header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && h.ok() && common_attrs(h, m, p, flist) && attr_string(h, "image_layout", "/frame_0/freq_i") &&
             attr_string(h, "diagnostics_layout", "/frame_0/diagnostics");
    }
    {
        Handle prm(H5Gcreate2(file, "/parameters", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        Handle cam(H5Gcreate2(prm, "camera", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && cam.ok() && attr_string(cam, "model", m.camera ? m.camera : "pinhole") &&
             attr_string(cam, "coordinate", m.coordinate ? m.coordinate : "bl") && attr_int(cam, "nx", nx) &&
             attr_int(cam, "ny", ny) && attr_double(cam, "radius", c.radius) &&
             attr_double(cam, "inclination_rad", c.inclination) && attr_double(cam, "fov", c.fov) &&
             attr_double(cam, "dsource", m.dsource_pc) && attr_double(cam, "mbh_solar", m.mbh_solar) &&
             attr_double(cam, "x_offset", c.x_offset) && attr_double(cam, "y_offset", c.y_offset) &&
             attr_string(cam, "grid_units", "dimensionless_slope") && attr_double(cam, "x_half_extent", xh) &&
             attr_double(cam, "y_half_extent", yh) && attr_double(cam, "dx", dx) && attr_double(cam, "dy", dy) &&
             attr_int(cam, "flux_scale_valid", scale_valid ? 1 : 0) &&
             attr_double(cam, "intensity_to_flux_jy_per_pixel", scale);
        Handle sp(H5Gcreate2(prm, "spacetime", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && sp.ok() && attr_string(sp, "metric", m.metric ? m.metric : "kerr") &&
             attr_string(sp, "coordinate", m.coordinate ? m.coordinate : "bl") && attr_double(sp, "spin", p.spin) &&
             attr_double(sp, "mass", p.mass) && attr_double(sp, "mbh_solar", m.mbh_solar);
        Handle in(H5Gcreate2(prm, "integration", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && in.ok() && attr_int(in, "adaptive", p.adaptive) && attr_double(in, "step", p.step) &&
             attr_double(in, "adaptive_tolerance", p.adaptive_tolerance) && attr_double(in, "min_step", p.min_step) &&
             attr_double(in, "max_step", p.max_step) && attr_int(in, "max_steps", p.max_steps) &&
             attr_double(in, "inner_radius", p.inner_radius) && attr_double(in, "outer_radius", p.outer_radius) &&
             attr_double(in, "max_radiation_step", p.max_radiation_step) &&
             attr_double(in, "max_radiation_depth", p.max_radiation_depth) &&
             attr_double(in, "max_absorption_depth", p.max_absorption_depth) &&
             attr_double(in, "max_faraday_depth", p.max_faraday_depth) &&
             attr_int(in, "direct_only", p.direct_only);
        Handle rad(H5Gcreate2(prm, "radiation", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && rad.ok() && attr_string(rad, "model", m.model ? m.model : "riaf") &&
             attr_int(rad, "faraday_rotation", p.emission_selection.faraday_rotation) &&
             attr_int(rad, "nfreq", m.nfreq) && attr_double(rad, "frequency_hz", m.frequency_hz[0]) &&
             attr_string(rad, "frequency_list_hz", flist) && attr_string(rad, "stokes_convention", "camera_frame") &&
             write_1d(rad, "frequencies", m.frequency_hz, (size_t)m.nfreq);
    }
    {
        std::vector<double> gx((size_t)nx), gy((size_t)ny);
        for (int i = 0; i < nx; ++i) gx[i] = (double(i) + 0.5) * dx - xh + c.x_offset;
        for (int j = 0; j < ny; ++j) gy[j] = (double(j) + 0.5) * dy - yh + c.y_offset;
        Handle g(H5Gcreate2(file, "/grid", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && g.ok() && attr_string(g, "units", "dimensionless_slope") && write_1d(g, "x", gx.data(), gx.size()) &&
             write_1d(g, "y", gy.data(), gy.size()) && write_1d(g, "frequency_hz", m.frequency_hz, (size_t)m.nfreq);
    }
    // packing per frequency (pack_transport_frequency_image,
    // tools/kpolaris_model_image.cpp:2659-2700)
    std::vector<double> sums(4 * (size_t)m.nfreq), fluxes(4 * (size_t)m.nfreq);
    int32_t returned = 0;
    for (size_t i = 0; i < np; ++i) returned += img->reason[i] == KP_REASON_REACHED_CAMERA;
    Handle frame(H5Gcreate2(file, "/frame_0", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    ok = ok && frame.ok() && attr_int(frame, "index", 0) && attr_int(frame, "nfreq", m.nfreq);
    std::vector<double> inv(4 * np), obs(4 * np);
    for (int f = 0; f < m.nfreq && ok; ++f) {
        const double nu = m.frequency_hz[f];
        const size_t off = (size_t)f * np;
        double flux[4];
        if (kp_pack_image(img->image_i + off, img->image_q + off, img->image_u + off, img->image_v + off, img->reason,
                          (int64_t)np, nu, qu_sign, scale, obs.data(), obs.data() + np, obs.data() + 2 * np,
                          obs.data() + 3 * np, flux) != KP_OK)
            return KP_INVALID_ARGUMENT;
        const double* src[4] = {img->image_i + off, img->image_q + off, img->image_u + off, img->image_v + off};
        const double sgn[4] = {1.0, double(qu_sign), double(qu_sign), 1.0};
        for (int s = 0; s < 4; ++s) {
            double sum = 0.0;
            for (size_t i = 0; i < np; ++i) {
                const double v = img->reason[i] == KP_REASON_REACHED_CAMERA ? sgn[s] * src[s][i] : 0.0;
                inv[s * np + i] = v;
                sum += v;
            }
            sums[(size_t)s * m.nfreq + f] = sum;
            fluxes[(size_t)s * m.nfreq + f] = flux[s];
        }
        const std::string gname = "freq_" + std::to_string(f);
        Handle fg(H5Gcreate2(frame, gname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && fg.ok() && attr_int(fg, "index", f) && attr_double(fg, "frequency", nu) &&
             attr_double(fg, "frequency_hz", nu) && attr_string(fg, "stokes_convention", "camera_frame") &&
             attr_string(fg, "polarization_basis", "camera_screen") && attr_string(fg, "evpa_0", evpa) &&
             attr_string(fg, "stokes_units", "cgs_specific_intensity") &&
             attr_double(fg, "intensity_to_flux_jy_per_pixel", scale) && attr_double(fg, "flux_I_Jy", flux[0]) &&
             attr_double(fg, "flux_Q_Jy", flux[1]) && attr_double(fg, "flux_U_Jy", flux[2]) &&
             attr_double(fg, "flux_V_Jy", flux[3]);
        const char* on[4] = {"I", "Q", "U", "V"};
        const char* in[4] = {"I_inv", "Q_inv", "U_inv", "V_inv"};
        for (int s = 0; s < 4 && ok; ++s)
            ok = write_2d(fg, on[s], obs.data() + s * np, nx, ny) && write_2d(fg, in[s], inv.data() + s * np, nx, ny);
    }
    {
        Handle dg(H5Gcreate2(frame, "diagnostics", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        std::vector<int32_t> total(np);
        for (size_t i = 0; i < np; ++i) total[i] = img->pass_a_steps[i] + img->steps[i];
        ok = ok && dg.ok() && write_2d_int(dg, "reason", img->reason, nx, ny) &&
             write_2d_int(dg, "pass_a_steps", img->pass_a_steps, nx, ny) && write_2d_int(dg, "steps", img->steps, nx, ny) &&
             write_2d_int(dg, "total_steps", total.data(), nx, ny);
        const char* rn[11] = {"closure_x", "closure_k", "final_null", "frame_error", "det_r", "overlap_r11",
                              "overlap_r12", "overlap_r21", "overlap_r22", "basis_identity_error",
                              "basis_rotation_angle"};
        for (int i = 0; i < 11 && ok; ++i) ok = write_2d(dg, rn[i], real[i], nx, ny);
    }
    {
        Handle r(H5Gcreate2(file, "/results", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        double maxcx = 0.0, maxck = 0.0, maxfe = 0.0, maxbe = 0.0;
        long long sa = 0, sb = 0;
        int32_t mxa = 0, mxb = 0, mxt = 0;
        for (size_t i = 0; i < np; ++i) {
            sa += img->pass_a_steps[i];
            sb += img->steps[i];
            mxa = std::max(mxa, img->pass_a_steps[i]);
            mxb = std::max(mxb, img->steps[i]);
            mxt = std::max(mxt, img->pass_a_steps[i] + img->steps[i]);
            if (img->reason[i] != KP_REASON_REACHED_CAMERA) continue;
            maxcx = std::max(maxcx, img->closure_x[i]);
            maxck = std::max(maxck, img->closure_k[i]);
            maxfe = std::max(maxfe, img->frame_error[i]);
            maxbe = std::max(maxbe, img->basis_identity_error[i]);
        }
        const size_t nf = (size_t)m.nfreq;
        ok = ok && r.ok() && attr_int(r, "pixels", (int32_t)np) && attr_int(r, "returned", returned) &&
             attr_int(r, "nfreq", m.nfreq) && attr_double(r, "max_closure_x", maxcx) &&
             attr_double(r, "max_closure_k", maxck) && attr_double(r, "max_frame_error", maxfe) &&
             attr_double(r, "max_basis_identity_error", maxbe) &&
             attr_double(r, "mean_pass_a_steps", double(sa) / double(np)) &&
             attr_double(r, "mean_pass_b_steps", double(sb) / double(np)) &&
             attr_double(r, "mean_total_steps", double(sa + sb) / double(np)) && attr_int(r, "max_pass_a_steps", mxa) &&
             attr_int(r, "max_pass_b_steps", mxb) && attr_int(r, "max_total_steps", mxt) &&
             attr_double(r, "sum_I_inv", sums[0]) && attr_double(r, "sum_Q_inv", sums[nf]) &&
             attr_double(r, "sum_U_inv", sums[2 * nf]) && attr_double(r, "sum_V_inv", sums[3 * nf]) &&
             attr_double(r, "intensity_to_flux_jy_per_pixel", scale) && attr_double(r, "flux_I_Jy", fluxes[0]) &&
             attr_double(r, "flux_Q_Jy", fluxes[nf]) && attr_double(r, "flux_U_Jy", fluxes[2 * nf]) &&
             attr_double(r, "flux_V_Jy", fluxes[3 * nf]) && write_1d(r, "frequency_hz", m.frequency_hz, nf) &&
             write_1d(r, "sum_I_inv_by_freq", sums.data(), nf) && write_1d(r, "sum_Q_inv_by_freq", sums.data() + nf, nf) &&
             write_1d(r, "sum_U_inv_by_freq", sums.data() + 2 * nf, nf) &&
             write_1d(r, "sum_V_inv_by_freq", sums.data() + 3 * nf, nf) &&
             write_1d(r, "flux_I_Jy_by_freq", fluxes.data(), nf) && write_1d(r, "flux_Q_Jy_by_freq", fluxes.data() + nf, nf) &&
             write_1d(r, "flux_U_Jy_by_freq", fluxes.data() + 2 * nf, nf) &&
             write_1d(r, "flux_V_Jy_by_freq", fluxes.data() + 3 * nf, nf);
    }
    if (!ok) return fail(KP_HIP_ERROR, std::string("HDF5 write of ") + path + " failed");
    return KP_OK;
}

// ---------------------------------------------------------------------------
// compare_stokes_hdf5.py (scripts/compare_stokes_hdf5.py:24-143)
// ---------------------------------------------------------------------------
namespace {

double finite_ratio(double num, double den) {
    if (den > 0.0) return num / den;
    return num == 0.0 ? 0.0 : std::numeric_limits<double>::infinity();
}

// normalize_evpa_zero / evpa_zero (scripts/kpolaris_evpa.py): "camera" is W,
// absent is W, root and frequency declarations must agree.
kp_status evpa_zero(hid_t root, hid_t group, std::string& out) {
    std::string a, b;
    const bool ha = read_attr_string(root, "evpa_0", a), hb = read_attr_string(group, "evpa_0", b);
    auto norm = [](std::string v, std::string& o) -> bool {
        if (v == "camera") v = "W";
        if (v != "N" && v != "W") return false;
        o = v;
        return true;
    };
    std::string na, nb;
    if (ha && !norm(a, na)) return fail(KP_INVALID_ARGUMENT, "unknown evpa_0=" + a);
    if (hb && !norm(b, nb)) return fail(KP_INVALID_ARGUMENT, "unknown evpa_0=" + b);
    if (ha && hb && na != nb) return fail(KP_INVALID_ARGUMENT, "conflicting evpa_0 declarations in the input file");
    out = ha ? na : (hb ? nb : "W");
    return KP_OK;
}

struct Loaded {
    std::vector<double> stokes[4];
    std::vector<hsize_t> shape;
    double frequency = 0.0;
    std::vector<double> gx, gy;
    bool has_grid = false;
};

kp_status load_image(const char* path, int frame, int fi, Loaded& L) {
    Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(KP_INVALID_ARGUMENT, std::string("cannot open ") + path);
    std::string schema;
    if (!read_attr_string(f, "schema", schema) || schema != "kpolaris_image")
        return fail(KP_INVALID_ARGUMENT, std::string(path) + ": expected schema='kpolaris_image'");
    const std::string gname = "/frame_" + std::to_string(frame) + "/freq_" + std::to_string(fi);
    if (!exists(f, gname)) return fail(KP_INVALID_ARGUMENT, std::string(path) + ": missing " + gname);
    Handle g(H5Gopen2(f, gname.c_str(), H5P_DEFAULT), H5Gclose);
    std::string ev;
    kp_status st = evpa_zero(f, g, ev);
    if (st != KP_OK) return st;
    const char* names[4] = {"I", "Q", "U", "V"};
    for (int s = 0; s < 4; ++s) {
        std::vector<hsize_t> dims;
        if (!read_array(g, names[s], H5T_NATIVE_DOUBLE, L.stokes[s], dims))
            return fail(KP_INVALID_ARGUMENT, std::string(path) + ": missing " + gname + "/" + names[s]);
        for (double v : L.stokes[s])
            if (!std::isfinite(v))
                return fail(KP_INVALID_ARGUMENT, std::string(path) + ": " + gname + "/" + names[s] + " contains NaN or Inf");
        if ((s == 1 || s == 2) && ev != "N")
            for (double& v : L.stokes[s]) v = -v;  // qu_basis_sign(source, "N")
        if (s == 0) L.shape = dims;
    }
    double fq = std::numeric_limits<double>::quiet_NaN();
    if (!read_attr_double(g, "frequency_hz", fq)) read_attr_double(f, "frequency_hz", fq);
    L.frequency = fq;
    std::vector<hsize_t> d1;
    L.has_grid = read_array(f, "/grid/x", H5T_NATIVE_DOUBLE, L.gx, d1) && read_array(f, "/grid/y", H5T_NATIVE_DOUBLE, L.gy, d1);
    return KP_OK;
}

}  // namespace

extern "C" kp_status kp_image_compare_hdf5(const char* reference, const char* candidate, int32_t frame,
                                           int32_t freq_index, kp_image_comparison* out) {
    kp_internal::clear_error();
    if (!reference || !candidate || !out) return fail(KP_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof(*out));
    QuietErrors quiet;
    Loaded R, G;
    kp_status st = load_image(reference, frame, freq_index, R);
    if (st != KP_OK) return st;
    st = load_image(candidate, frame, freq_index, G);
    if (st != KP_OK) return st;
    if (R.has_grid != G.has_grid) return fail(KP_INVALID_ARGUMENT, "grid mismatch: only one file contains the grid");
    if (R.has_grid) {
        const std::vector<double>* ra[2] = {&R.gx, &R.gy};
        const std::vector<double>* ga[2] = {&G.gx, &G.gy};
        for (int k = 0; k < 2; ++k) {
            if (ra[k]->size() != ga[k]->size()) return fail(KP_INVALID_ARGUMENT, "grid mismatch: sizes differ");
            double mx = 0.0, scale = 0.0;
            for (size_t i = 0; i < ra[k]->size(); ++i) {
                mx = std::max(mx, std::abs((*ra[k])[i] - (*ga[k])[i]));
                scale = std::max({scale, std::abs((*ra[k])[i]), std::abs((*ga[k])[i])});
            }
            out->grid_max_abs_difference[k] = mx;
            if (mx > 64.0 * std::numeric_limits<double>::epsilon() * scale)
                return fail(KP_INVALID_ARGUMENT, k == 0 ? "grid mismatch in grid_x" : "grid mismatch in grid_y");
        }
    }
    if (R.shape != G.shape) return fail(KP_INVALID_ARGUMENT, "image shape mismatch");
    const double fr = R.frequency, fg = G.frequency;
    if (!(std::abs(fr - fg) <= 1e-12 * std::max(std::abs(fr), std::abs(fg))))
        return fail(KP_INVALID_ARGUMENT, "frequency mismatch");
    out->frequency_hz = fr;
    out->ny = R.shape.size() > 0 ? (int32_t)R.shape[0] : 0;
    out->nx = R.shape.size() > 1 ? (int32_t)R.shape[1] : 0;
    for (int s = 0; s < 4; ++s) {
        const auto& r = R.stokes[s];
        const auto& g = G.stokes[s];
        double rs = 0.0, gs = 0.0, l1 = 0.0, r1 = 0.0, d2 = 0.0, r2 = 0.0, dmax = 0.0, rmax = 0.0;
        for (size_t i = 0; i < r.size(); ++i) {
            const double d = g[i] - r[i];
            rs += r[i];
            gs += g[i];
            l1 += std::abs(d);
            r1 += std::abs(r[i]);
            d2 += d * d;
            r2 += r[i] * r[i];
            dmax = std::max(dmax, std::abs(d));
            rmax = std::max(rmax, std::abs(r[i]));
        }
        kp_stokes_norms& n = out->stokes[s];
        n.reference_sum = rs;
        n.candidate_sum = gs;
        n.sum_relative = rs != 0.0 ? (gs - rs) / rs : (gs == 0.0 ? 0.0 : std::numeric_limits<double>::infinity());
        n.l1_relative = finite_ratio(l1, r1);
        n.rms_relative = std::sqrt(finite_ratio(d2, r2));
        n.linf_over_peak = finite_ratio(dmax, rmax);
        out->worst_relative_l1 = std::max(out->worst_relative_l1, n.l1_relative);
        out->worst_linf_over_peak = std::max(out->worst_linf_over_peak, n.linf_over_peak);
    }
    // diagnostics (compare_diagnostics, scripts/compare_stokes_hdf5.py:96-143)
    Handle fr_(H5Fopen(reference, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    Handle fc_(H5Fopen(candidate, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    const std::string dg = "/frame_" + std::to_string(frame) + "/diagnostics/";
    const char* dn[4] = {"reason", "pass_a_steps", "steps", "total_steps"};
    for (int k = 0; k < 4; ++k) {
        std::vector<int64_t> a, b;
        std::vector<hsize_t> da, db;
        if (!read_array(fr_, dg + dn[k], H5T_NATIVE_INT64, a, da) || !read_array(fc_, dg + dn[k], H5T_NATIVE_INT64, b, db))
            return fail(KP_INVALID_ARGUMENT, std::string("missing diagnostic ") + dn[k]);
        if (da != db) return fail(KP_INVALID_ARGUMENT, std::string("diagnostic ") + dn[k] + ": shape mismatch");
        for (size_t i = 0; i < a.size(); ++i) {
            const int64_t d = std::llabs(b[i] - a[i]);
            out->discrete_mismatched[k] += d != 0;
            out->discrete_max_abs_difference[k] = std::max(out->discrete_max_abs_difference[k], d);
        }
    }
    const char* cn[5] = {"closure_x", "closure_k", "final_null", "frame_error", "basis_identity_error"};
    for (int k = 0; k < 5; ++k) {
        std::vector<double> a, b;
        std::vector<hsize_t> da, db;
        if (!read_array(fr_, dg + cn[k], H5T_NATIVE_DOUBLE, a, da) || !read_array(fc_, dg + cn[k], H5T_NATIVE_DOUBLE, b, db))
            return fail(KP_INVALID_ARGUMENT, std::string("missing diagnostic ") + cn[k]);
        if (da != db) return fail(KP_INVALID_ARGUMENT, std::string("diagnostic ") + cn[k] + ": shape mismatch");
        double l1 = 0.0, r1 = 0.0, mx = 0.0;
        for (size_t i = 0; i < a.size(); ++i) {
            if (!std::isfinite(a[i]) || !std::isfinite(b[i]))
                return fail(KP_INVALID_ARGUMENT, std::string("diagnostic ") + cn[k] + " contains NaN or Inf");
            l1 += std::abs(b[i] - a[i]);
            r1 += std::abs(a[i]);
            mx = std::max(mx, std::abs(b[i] - a[i]));
        }
        out->continuous_l1_relative[k] = finite_ratio(l1, r1);
        out->continuous_linf[k] = mx;
    }
    return KP_OK;
}
