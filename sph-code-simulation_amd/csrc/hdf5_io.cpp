// hdf5_io.cpp -- HDF5 snapshot files in the reference's layout
// (HDF5Writer, src/writers/hdf5_writer.cpp:20-454), built as its own library
// against the HDF5 C API so that the core library needs no HDF5 runtime.
// The metadata JSON comes from libsphgpu (sphgpu_output_meta_json /
// sphgpu_output_meta_from_json), the device state from sphgpu_download /
// sphgpu_upload.
#include <hdf5.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sphgpu_hdf5.h"

namespace {

// the datasets of /particles in the reference's order (hdf5_writer.cpp:
// 207-227); row -1 marks an axis the block does not have
struct Column {
    const char *name;
    int row;
    bool integer;
};

std::vector<Column> particle_columns(int dim)
{
    static const char *const vec_names[3][3] = {
        {"pos_x", "pos_y", "pos_z"}, {"vel_x", "vel_y", "vel_z"}, {"acc_x", "acc_y", "acc_z"}};
    const int vec_rows[3] = {SPHGPU_F_POS, SPHGPU_F_VEL, SPHGPU_F_ACC};
    std::vector<Column> c;
    c.push_back({"id", SPHGPU_F_ID, true});
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) c.push_back({vec_names[v][a], a < dim ? vec_rows[v] + a : -1, false});
    const std::pair<const char *, int> rest[] = {
        {"mass", SPHGPU_F_MASS}, {"dens", SPHGPU_F_DENS}, {"pres", SPHGPU_F_PRES}, {"ene", SPHGPU_F_ENE},
        {"sml", SPHGPU_F_SML}, {"sound", SPHGPU_F_SOUND}, {"alpha", SPHGPU_F_ALPHA},
        {"balsara", SPHGPU_F_BALSARA}, {"gradh", SPHGPU_F_GRADH}, {"phi", SPHGPU_F_PHI}};
    for (const auto &r : rest) c.push_back({r.first, r.second, false});
    c.push_back({"neighbor", SPHGPU_F_NEIGHBOR, true});
    return c;
}

bool attr_scalar(hid_t loc, const char *name, hid_t type, const void *v)
{
    const hsize_t dims[1] = {1};
    const hid_t sp = H5Screate_simple(1, dims, nullptr);
    const hid_t a = H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, type, v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

// HDF5Writer::write_string_attribute (hdf5_writer.cpp:273-293): a scalar
// fixed-length NUL-terminated C string of size length + 1
bool attr_string(hid_t loc, const char *name, const std::string &v)
{
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, v.size() + 1);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    const hid_t a = H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, t, v.c_str()) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Tclose(t);
    H5Sclose(sp);
    return ok;
}

bool read_attr_string(hid_t loc, const char *name, std::string &v)
{
    if (H5Aexists(loc, name) <= 0) return false;
    const hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    if (a < 0) return false;
    const hid_t t = H5Aget_type(a);
    const size_t sz = H5Tget_size(t);
    std::vector<char> buf(sz + 1, '\0');
    const bool ok = H5Aread(a, t, buf.data()) >= 0;
    v = std::string(buf.data());
    H5Tclose(t);
    H5Aclose(a);
    return ok;
}

std::string meta_text(const sphgpu_output_meta &m, int which)
{
    int32_t need = 0;
    sphgpu_output_meta_json(&m, which, nullptr, 0, &need);
    std::vector<char> buf((size_t)std::max(need, 1));
    if (sphgpu_output_meta_json(&m, which, buf.data(), need, nullptr)) return std::string();
    return std::string(buf.data());
}

const char *sph_name(int t)
{
    switch (t) {
    case SPHGPU_SSPH: return "Standard SPH";
    case SPHGPU_DISPH: return "Density Independent SPH";
    case SPHGPU_GSPH: return "Godunov SPH";
    default: return "Unknown";
    }
}

const char *kernel_type_name(int k)
{
    switch (k) {
    case SPHGPU_CUBIC_SPLINE: return "Cubic Spline";
    case SPHGPU_WENDLAND: return "Wendland C4";
    default: return "Unknown";
    }
}

// write_metadata_group (hdf5_writer.cpp:61-126)
bool write_metadata(hid_t file, const sphgpu_output_meta &m)
{
    const hid_t g = H5Gcreate2(file, "/metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return false;
    bool ok = attr_scalar(g, "format_version", H5T_NATIVE_INT, &m.format_version) &&
              attr_scalar(g, "step", H5T_NATIVE_INT, &m.step) &&
              attr_scalar(g, "particle_count", H5T_NATIVE_INT, &m.particle_count) &&
              attr_scalar(g, "time_code", H5T_NATIVE_DOUBLE, &m.time_code) &&
              attr_scalar(g, "time_physical", H5T_NATIVE_DOUBLE, &m.time_physical) &&
              attr_scalar(g, "gamma", H5T_NATIVE_DOUBLE, &m.gamma);
    ok = ok && attr_string(g, "simulation_name", m.simulation_name) && attr_string(g, "timestamp", m.timestamp) &&
         attr_string(g, "sph_type", sph_name(m.sph_type)) &&
         attr_string(g, "kernel_type", kernel_type_name(m.kernel)) &&
         attr_string(g, "unit_system", meta_text(m, 1)) && attr_string(g, "metadata_json", meta_text(m, 0));
    H5Gclose(g);
    return ok;
}

// write_energy_group (hdf5_writer.cpp:237-271)
bool write_energy(hid_t file, const sphgpu_output_meta &m)
{
    const hid_t g = H5Gcreate2(file, "/energy", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return false;
    const hsize_t dims[1] = {1};
    const hid_t sp = H5Screate_simple(1, dims, nullptr);
    bool ok = true;
    const std::pair<const char *, const double *> e[] = {{"kinetic", &m.kinetic_energy},
                                                         {"thermal", &m.thermal_energy},
                                                         {"potential", &m.potential_energy},
                                                         {"total", &m.total_energy}};
    for (const auto &x : e) {
        const hid_t d = H5Dcreate2(g, x.first, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        ok = ok && d >= 0 && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, x.second) >= 0;
        if (d >= 0) H5Dclose(d);
    }
    H5Sclose(sp);
    H5Gclose(g);
    return ok;
}

// write_particles_group (hdf5_writer.cpp:128-235)
bool write_particles(hid_t file, int dim, int32_t n, const double *f, int level)
{
    const hid_t g = H5Gcreate2(file, "/particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return false;
    const hsize_t N = (hsize_t)n;
    hid_t plist = H5P_DEFAULT;
    if (level > 0) {
        plist = H5Pcreate(H5P_DATASET_CREATE);
        const hsize_t chunk[1] = {N < 1024 ? N : 1024};
        H5Pset_chunk(plist, 1, chunk);
        H5Pset_deflate(plist, (unsigned)level);
    }
    bool ok = true;
    std::vector<double> col((size_t)n);
    std::vector<int> icol((size_t)n);
    for (const Column &c : particle_columns(dim)) {
        const hsize_t dims[1] = {N};
        const hid_t sp = H5Screate_simple(1, dims, nullptr);
        const hid_t type = c.integer ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE;
        const hid_t d = H5Dcreate2(g, c.name, type, sp, H5P_DEFAULT, plist, H5P_DEFAULT);
        if (c.integer) {
            for (int32_t i = 0; i < n; ++i) icol[(size_t)i] = (int)f[(size_t)c.row * (size_t)n + (size_t)i];
            ok = ok && d >= 0 && H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, icol.data()) >= 0;
        } else {
            for (int32_t i = 0; i < n; ++i)
                col[(size_t)i] = c.row < 0 ? 0.0 : f[(size_t)c.row * (size_t)n + (size_t)i];
            ok = ok && d >= 0 && H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, col.data()) >= 0;
        }
        if (d >= 0) H5Dclose(d);
        H5Sclose(sp);
    }
    if (plist != H5P_DEFAULT) H5Pclose(plist);
    H5Gclose(g);
    return ok;
}

// HDF5's own error stack printing is off: failures come back as codes
struct QuietErrors {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    QuietErrors()
    {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietErrors() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

int row_count(hid_t file, hsize_t &N)
{
    const hid_t g = H5Gopen2(file, "/particles", H5P_DEFAULT);
    if (g < 0) return SPHGPU_E_ARG;
    const hid_t d = H5Dopen2(g, "id", H5P_DEFAULT);
    if (d < 0) {
        H5Gclose(g);
        return SPHGPU_E_ARG;
    }
    const hid_t sp = H5Dget_space(d);
    hsize_t dims[1] = {0};
    const int nd = H5Sget_simple_extent_ndims(sp);
    if (nd == 1) H5Sget_simple_extent_dims(sp, dims, nullptr);
    H5Sclose(sp);
    H5Dclose(d);
    H5Gclose(g);
    if (nd != 1) return SPHGPU_E_ARG;
    N = dims[0];
    return 0;
}

} // namespace

extern "C" {

int sphgpu_write_hdf5_fields(const char *path, const sphgpu_output_meta *meta, int32_t dim, int32_t n,
                             const double *f, int32_t compression)
{
    if (!path || !meta || n < 0 || (n > 0 && !f) || dim < 1 || dim > 3) return SPHGPU_E_ARG;
    const int level = std::min(9, std::max(0, (int)compression));  // HDF5Writer's clamp (:7-14)
    QuietErrors quiet;
    const hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return SPHGPU_E_ARG;
    bool ok = write_metadata(file, *meta) && write_energy(file, *meta);
    ok = ok && n > 0 && write_particles(file, dim, n, f, level);
    ok = (H5Fclose(file) >= 0) && ok;
    return ok ? 0 : SPHGPU_E_ARG;
}

int sphgpu_write_snapshot_hdf5(sphgpu_ctx *ctx, const char *path, const sphgpu_output_meta *meta,
                               int32_t compression)
{
    if (!ctx || !path || !meta) return SPHGPU_E_ARG;
    int32_t n = 0;
    sphgpu_params p;
    int rc;
    if ((rc = sphgpu_get_params(ctx, &p, &n))) return rc;
    std::vector<double> f((size_t)SPHGPU_NFIELDS * n);
    if ((rc = sphgpu_download(ctx, f.data(), SPHGPU_MASK_ALL))) return rc;
    return sphgpu_write_hdf5_fields(path, meta, p.dim, n, f.data(), compression);
}

int sphgpu_hdf5_row_count(const char *path, int32_t *rows)
{
    if (!path || !rows) return SPHGPU_E_ARG;
    QuietErrors quiet;
    const hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return SPHGPU_E_ARG;
    hsize_t N = 0;
    const int rc = row_count(file, N);
    H5Fclose(file);
    if (rc) return rc;
    *rows = (int32_t)N;
    return 0;
}

int sphgpu_read_hdf5_fields(const char *path, sphgpu_output_meta *meta, int32_t n, double *f)
{
    if (!path || n < 0 || (n > 0 && !f)) return SPHGPU_E_ARG;
    QuietErrors quiet;
    const hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return SPHGPU_E_ARG;
    hsize_t N = 0;
    int rc = row_count(file, N);
    if (!rc && N != (hsize_t)n) rc = SPHGPU_E_ARG;
    std::vector<double> block;
    if (!rc) {
        // read_particles (hdf5_writer.cpp:339-423): a dataset that does not
        // open leaves its column at 0; rows are placed by id
        const hid_t g = H5Gopen2(file, "/particles", H5P_DEFAULT);
        block.assign((size_t)SPHGPU_NFIELDS * (size_t)n, 0.0);
        std::vector<double> col((size_t)n);
        std::vector<int> icol((size_t)n), ids((size_t)n);
        bool ok = true;
        for (const Column &c : particle_columns(3)) {
            const hid_t d = H5Dopen2(g, c.name, H5P_DEFAULT);
            std::fill(col.begin(), col.end(), 0.0);
            std::fill(icol.begin(), icol.end(), 0);
            if (d >= 0) {
                if (c.integer) ok = ok && H5Dread(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, icol.data()) >= 0;
                else ok = ok && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, col.data()) >= 0;
                H5Dclose(d);
            }
            if (c.row == SPHGPU_F_ID) ids = icol;
            for (int32_t i = 0; i < n; ++i) col[(size_t)i] = c.integer ? (double)icol[(size_t)i] : col[(size_t)i];
            for (int32_t i = 0; i < n; ++i) {
                const int id = ids[(size_t)i];
                if (id < 0 || id >= n) {
                    ok = false;
                    break;
                }
                block[(size_t)c.row * (size_t)n + (size_t)id] = col[(size_t)i];
            }
        }
        H5Gclose(g);
        if (!ok) rc = SPHGPU_E_ARG;
    }
    sphgpu_output_meta m;
    std::memset(&m, 0, sizeof(m));
    if (!rc) {
        // read_metadata (hdf5_writer.cpp:295-314): metadata_json through
        // from_json; without it the metadata keeps OutputMetadata()'s values
        const hid_t g = H5Gopen2(file, "/metadata", H5P_DEFAULT);
        if (g < 0) {
            rc = SPHGPU_E_ARG;
        } else {
            std::string js;
            rc = sphgpu_output_meta_from_json(read_attr_string(g, "metadata_json", js) ? js.c_str() : "{}", &m);
            H5Gclose(g);
        }
    }
    H5Fclose(file);
    if (rc) return rc;
    if (n > 0) std::memcpy(f, block.data(), block.size() * sizeof(double));
    if (meta) *meta = m;
    return 0;
}

int sphgpu_read_snapshot_hdf5(sphgpu_ctx *ctx, const char *path, sphgpu_output_meta *meta)
{
    if (!ctx || !path) return SPHGPU_E_ARG;
    int32_t n = 0;
    sphgpu_params p;
    int rc;
    if ((rc = sphgpu_get_params(ctx, &p, &n))) return rc;
    std::vector<double> f((size_t)SPHGPU_NFIELDS * n);
    sphgpu_output_meta m;
    if ((rc = sphgpu_read_hdf5_fields(path, &m, n, f.data()))) return rc;
    if ((rc = sphgpu_upload(ctx, f.data(), SPHGPU_MASK_ALL))) return rc;
    if ((rc = sphgpu_set_time(ctx, m.time_code))) return rc;
    if (meta) *meta = m;
    return 0;
}

} // extern "C"
