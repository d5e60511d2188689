// HDF5 field-dump writer in the reference's on-disk layout (include/wipems_dump.h).
//
// The file stays open between samples (the reference reopens it for every
// dataset; the resulting objects are the same).  Groups are created on first
// use, like HDF5_File_Writer::OpenGroup (tools/hdf5_file_writer.cpp:42-88).
#include "wipems_dump.h"

#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>
#include <cmath>
#include <cfloat>

struct wdump_file {
    hid_t file = -1;
    uint32_t n[3] = {0, 0, 0};
    std::vector<float> buf;
};

namespace {

thread_local std::string g_err;

int fail(const char* fmt, ...) {
    char m[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(m, sizeof m, fmt, ap);
    va_end(ap);
    g_err = m;
    return -1;
}

// opens (creating where missing) every group on a slash-separated path
hid_t open_group(hid_t file, const std::string& path) {
    hid_t g = H5Gopen2(file, "/", H5P_DEFAULT);
    size_t at = 0;
    while (g >= 0 && at < path.size()) {
        size_t end = path.find('/', at);
        if (end == std::string::npos) end = path.size();
        const std::string part = path.substr(at, end - at);
        at = end + 1;
        if (part.empty()) continue;
        hid_t next = H5Lexists(g, part.c_str(), H5P_DEFAULT) > 0 ? H5Gopen2(g, part.c_str(), H5P_DEFAULT)
                                                                   : H5Gcreate2(g, part.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H5Gclose(g);
        g = next;
    }
    return g;
}

int write_attr(hid_t loc, const char* name, hid_t type, const void* value, hsize_t count) {
    hid_t sp = H5Screate_simple(1, &count, nullptr);
    hid_t a = H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, type, value) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok ? 0 : fail("writing attribute %s failed", name);
}

// dataset of rank `rank` and extents `dims` holding `data` as float
int write_float_set(hid_t grp, const std::string& name, int rank, const hsize_t* dims, const float* data, hid_t* keep) {
    hid_t sp = H5Screate_simple(rank, dims, nullptr);
    hid_t ds = H5Dcreate2(grp, name.c_str(), H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = ds >= 0 && H5Dwrite(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    H5Sclose(sp);
    if (!ok) {
        if (ds >= 0) H5Dclose(ds);
        return fail("writing dataset %s failed", name.c_str());
    }
    if (keep) *keep = ds;
    else H5Dclose(ds);
    return 0;
}

// a value rounded to float as the reference stores dump fields: on a thread
// with flush-to-zero set (Engine_sse's constructor calls Denormal::Disable,
// FDTD/engine_sse.cpp:38-39), so a float denormal becomes a zero of its sign
inline float to_float_ftz(double v) {
    const float f = (float)v;
    return std::fabs(f) < FLT_MIN ? std::copysign(0.0f, f) : f;
}

// [3][nx][ny][nz] (z fastest) -> [3][nz][ny][nx] (x fastest), the order of
// HDF5_File_Writer::WriteVectorField's buffer fill
template <typename F>
void transpose(const uint32_t n[3], F get, std::vector<float>& out) {
    const size_t nx = n[0], ny = n[1], nz = n[2];
    out.resize(3 * nx * ny * nz);
    size_t pos = 0;
    for (int c = 0; c < 3; ++c)
        for (size_t k = 0; k < nz; ++k)
            for (size_t j = 0; j < ny; ++j)
                for (size_t i = 0; i < nx; ++i) out[pos++] = to_float_ftz(get(((c * nx + i) * ny + j) * nz + k));
}

}  // namespace

extern "C" {

int wdump_last_error(char* buf, size_t len) {
    if (!buf || !len) return -1;
    snprintf(buf, len, "%s", g_err.c_str());
    return 0;
}

int wdump_create(const char* path, const double* lines_x, uint32_t nx, const double* lines_y, uint32_t ny,
                 const double* lines_z, uint32_t nz, double scaling, wdump_file** out) {
    if (!path || !out || !lines_x || !lines_y || !lines_z || !nx || !ny || !nz) return fail("bad argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    wdump_file* w = new wdump_file();
    w->n[0] = nx; w->n[1] = ny; w->n[2] = nz;
    w->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (w->file < 0) {
        delete w;
        return fail("creating %s failed", path);
    }
    hid_t mesh = open_group(w->file, "/Mesh");
    if (mesh < 0) {
        wdump_close(w);
        return fail("creating /Mesh failed");
    }
    const double* lines[3] = {lines_x, lines_y, lines_z};
    const char* names[3] = {"x", "y", "z"};
    const float fs = (float)scaling;
    for (int a = 0; a < 3; ++a) {
        std::vector<float> v(w->n[a]);
        for (uint32_t i = 0; i < w->n[a]; ++i) v[i] = (float)lines[a][i] * fs;   // float product, as WriteRectMesh
        const hsize_t d = w->n[a];
        if (write_float_set(mesh, names[a], 1, &d, v.data(), nullptr)) {
            H5Gclose(mesh);
            wdump_close(w);
            return -1;
        }
    }
    H5Gclose(mesh);
    hid_t root = H5Gopen2(w->file, "/", H5P_DEFAULT);
    const double version = 0.2;
    const int rc = write_attr(root, "openEMS_HDF5_version", H5T_NATIVE_DOUBLE, &version, 1);
    H5Gclose(root);
    if (rc) {
        wdump_close(w);
        return -1;
    }
    *out = w;
    return 0;
}

int wdump_write_td(wdump_file* w, uint32_t ts, double time, const float* field) {
    if (!w || w->file < 0 || !field) return fail("bad argument");
    hid_t g = open_group(w->file, "/FieldData/TD");
    if (g < 0) return fail("opening /FieldData/TD failed");
    char name[32];
    snprintf(name, sizeof name, "%08u", ts);
    transpose(w->n, [&](size_t q) { return field[q]; }, w->buf);
    const hsize_t dims[4] = {3, w->n[2], w->n[1], w->n[0]};
    hid_t ds = -1;
    int rc = write_float_set(g, name, 4, dims, w->buf.data(), &ds);
    if (!rc) {
        const float t = (float)time;
        rc = write_attr(ds, "time", H5T_NATIVE_FLOAT, &t, 1);
        H5Dclose(ds);
    }
    H5Gclose(g);
    return rc;
}

int wdump_write_fd(wdump_file* w, uint32_t n_freq, const double* freqs, const double* acc) {
    if (!w || w->file < 0 || (n_freq && (!freqs || !acc))) return fail("bad argument");
    hid_t g = open_group(w->file, "/FieldData/FD");
    if (g < 0) return fail("opening /FieldData/FD failed");
    int rc = write_attr(g, "frequency", H5T_NATIVE_DOUBLE, freqs, n_freq);
    const size_t per = 3 * (size_t)w->n[0] * w->n[1] * w->n[2];
    const hsize_t dims[4] = {3, w->n[2], w->n[1], w->n[0]};
    for (uint32_t k = 0; k < n_freq && !rc; ++k) {
        const double* a = acc + 2 * per * k;
        const float f = (float)freqs[k];
        for (int part = 0; part < 2 && !rc; ++part) {
            transpose(w->n, [&](size_t q) { return a[2 * q + part]; }, w->buf);
            char name[40];
            snprintf(name, sizeof name, "f%u_%s", k, part ? "imag" : "real");
            hid_t ds = -1;
            rc = write_float_set(g, name, 4, dims, w->buf.data(), &ds);
            if (!rc) {
                rc = write_attr(ds, "frequency", H5T_NATIVE_FLOAT, &f, 1);
                H5Dclose(ds);
            }
        }
    }
    H5Gclose(g);
    return rc;
}

int wdump_write_sar(wdump_file* w, uint32_t n_freq, const double* freqs, const double* acc, const float* kappa, const float* density,
                    const float* volume, int raw) {
    if (!w || w->file < 0 || !kappa || !density || !volume || (n_freq && (!freqs || !acc))) return fail("bad argument");
    hid_t g = open_group(w->file, "/FieldData/FD");
    if (g < 0) return fail("opening /FieldData/FD failed");
    int rc = write_attr(g, "frequency", H5T_NATIVE_DOUBLE, freqs, n_freq);
    const size_t nx = w->n[0], ny = w->n[1], nz = w->n[2], cells = nx * ny * nz, per = 3 * cells;
    const hsize_t dims3[3] = {nz, ny, nx};
    const hsize_t dims4[4] = {3, nz, ny, nx};
    // scalar [nx][ny][nz] -> [nz][ny][nx] (HDF5_File_Writer::WriteScalarField, tools/hdf5_file_writer.cpp:198-213)
    auto scalar = [&](const std::string& name, hid_t grp, auto get) -> hid_t {
        w->buf.resize(cells);
        size_t pos = 0;
        for (size_t k = 0; k < nz; ++k)
            for (size_t j = 0; j < ny; ++j)
                for (size_t i = 0; i < nx; ++i) w->buf[pos++] = to_float_ftz(get((i * ny + j) * nz + k));
        hid_t ds = -1;
        rc = write_float_set(grp, name, 3, dims3, w->buf.data(), &ds);
        return rc ? -1 : ds;
    };
    for (uint32_t f = 0; f < n_freq && !rc; ++f) {
        const double* a = acc + 2 * per * f;
        char name[40];
        snprintf(name, sizeof name, "f%u", f);
        if (raw) {
            // SAR_RAW_DATA: the E accumulators (Common/processfields_sar.cpp:245-256)
            for (int part = 0; part < 2 && !rc; ++part) {
                transpose(w->n, [&](size_t q) { return a[2 * q + part]; }, w->buf);
                rc = write_float_set(g, std::string(name) + (part ? "_imag" : "_real"), 4, dims4, w->buf.data(), nullptr);
            }
            continue;
        }
        // local SAR (SAR_Calculation::CalcLocalSAR / CalcLocalPowerDensity with cell
        // conductivity, tools/sar_calculation.cpp:162-230): per cell
        // P = 0.5 * sum_n kappa |E_n| |E_n|, SAR = P / density (0 where density is 0),
        // total power sum(P * volume) in the reference's loop order
        std::vector<double> sar(cells);
        double power = 0;
        for (size_t q = 0; q < cells; ++q) {
            double l = 0;
            for (int n = 0; n < 3; ++n) {
                const double* e = a + 2 * (n * cells + q);
                const double m = std::hypot(e[0], e[1]);
                l += (double)kappa[q] * m * m;
            }
            l *= 0.5;
            power += l * (double)volume[q];
            sar[q] = density[q] > 0 ? l / (double)density[q] : 0.0;
        }
        hid_t ds = scalar(name, g, [&](size_t q) { return sar[q]; });
        if (ds >= 0) {
            const float fr = (float)freqs[f], pw = to_float_ftz(power);
            rc = write_attr(ds, "frequency", H5T_NATIVE_FLOAT, &fr, 1);
            if (!rc) rc = write_attr(ds, "power", H5T_NATIVE_FLOAT, &pw, 1);
            H5Dclose(ds);
        }
    }
    H5Gclose(g);
    if (raw && !rc) {
        hid_t c = open_group(w->file, "/CellData");
        if (c < 0) return fail("opening /CellData failed");
        const float* cell[3] = {kappa, density, volume};
        const char* names[3] = {"Conductivity", "Density", "Volume"};
        for (int t = 0; t < 3 && !rc; ++t) {
            hid_t ds = scalar(names[t], c, [&](size_t q) { return (double)cell[t][q]; });
            if (ds >= 0) H5Dclose(ds);
        }
        H5Gclose(c);
    }
    return rc;
}

int wdump_close(wdump_file* w) {
    if (!w) return fail("null handle");
    int rc = 0;
    if (w->file >= 0 && H5Fclose(w->file) < 0) rc = fail("closing the file failed");
    delete w;
    return rc;
}

}  // extern "C"
