// io_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the CPU checker of the VTK and
// HDF5 snapshot formats and of the metadata JSON; nothing in the product
// links it).
//
// A restatement, with the reference's own machinery, of:
//   OutputMetadata::to_json / from_json (src/output_metadata.cpp:67-168) and
//       UnitSystem::to_json / from_json (src/units.cpp:206-250), built and
//       printed by nlohmann::json.  The reference pins nlohmann_json 3.11.3
//       (CMakeLists.txt:18-27, fetched, absent here); this uses the single
//       header of version 3.1.1 in the image (/opt/conda/include/json.hpp),
//       whose dump(2) and double printing (to_chars: Grisu2 with the same
//       cached powers, boundaries and "%.17g"-free formatting) are the
//       algorithm the later version keeps;
//   VTKWriter (src/writers/vtk_writer.cpp:17-331): std::ofstream, operator<<
//       for ASCII, byte-swapped floats and ints for BINARY;
//   HDF5Writer (src/writers/hdf5_writer.cpp:20-454): the HDF5 C API calls of
//       the writer and of read_particles / read_metadata.
// orc_hdf5_describe walks a file with the HDF5 API and prints its groups,
// attributes and datasets (types, extents, chunking, filters, values), so
// two files can be compared object by object (HDF5 files carry object
// timestamps, so their bytes differ between runs).
#include <hdf5.h>

#include "/opt/conda/include/json.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../include/sphgpu.h"

namespace {

using json = nlohmann::json;

const char *type_name(int t)
{
    switch (t) {
    case SPHGPU_SSPH: return "Standard SPH";
    case SPHGPU_DISPH: return "Density Independent SPH";
    case SPHGPU_GSPH: return "Godunov SPH";
    default: return "Unknown";
    }
}

const char *kern_name(int k)
{
    switch (k) {
    case SPHGPU_CUBIC_SPLINE: return "Cubic Spline";
    case SPHGPU_WENDLAND: return "Wendland C4";
    default: return "Unknown";
    }
}

// UnitSystem as the metadata carries it: the type from its name, the names
// from the type (src/units.cpp:125-203), the derived factors by
// compute_derived_factors (src/units.cpp:55-60)
int unit_enum(const std::string &t)
{
    if (t == "Code") return 0;
    if (t == "Galactic") return 1;
    if (t == "SI") return 2;
    if (t == "CGS") return 3;
    return 4;
}

json units_json(const sphgpu_output_meta &m)
{
    static const char *dens_name[5] = {"code_density", "M_sun/kpc^3", "kg/m^3", "g/cm^3", "unknown"};
    static const char *pres_name[5] = {"code_pressure", "erg/cm^3", "Pa", "dyne/cm^2", "unknown"};
    const int t = unit_enum(m.unit_type);
    const double L = m.length_to_cgs, M = m.mass_to_cgs, V = m.velocity_to_cgs;
    const double energy = M * V * V;
    const double density = M / (L * L * L);
    const double pressure = energy / (L * L * L);
    json j;
    j["type"] = std::string(m.unit_type);
    j["type_enum"] = t;
    j["length_unit"] = std::string(m.length_unit);
    j["mass_unit"] = std::string(m.mass_unit);
    j["time_unit"] = std::string(m.time_unit);
    j["velocity_unit"] = std::string(m.velocity_unit);
    j["energy_unit"] = std::string(m.energy_unit);
    j["density_unit"] = dens_name[t];
    j["pressure_unit"] = pres_name[t];
    j["length_to_cgs"] = L;
    j["mass_to_cgs"] = M;
    j["time_to_cgs"] = m.time_to_cgs;
    j["velocity_to_cgs"] = V;
    j["energy_to_cgs"] = energy;
    j["density_to_cgs"] = density;
    j["pressure_to_cgs"] = pressure;
    return j;
}

json meta_json(const sphgpu_output_meta &m)
{
    json j;
    j["simulation_name"] = std::string(m.simulation_name);
    j["timestamp"] = std::string(m.timestamp);
    j["format_version"] = m.format_version;
    j["step"] = m.step;
    j["time_code"] = m.time_code;
    j["time_physical"] = m.time_physical;
    j["particle_count"] = m.particle_count;
    j["sph_type"] = type_name(m.sph_type);
    j["sph_type_enum"] = m.sph_type;
    j["kernel_type"] = kern_name(m.kernel);
    j["kernel_type_enum"] = m.kernel;
    j["gamma"] = m.gamma;
    j["gravitational_constant"] = m.gravitational_constant;
    j["neighbor_number"] = m.neighbor_number;
    j["use_gravity"] = m.use_gravity != 0;
    j["use_balsara"] = m.use_balsara != 0;
    j["use_time_dependent_av"] = m.use_time_dependent_av != 0;
    j["units"] = units_json(m);
    j["is_relaxation"] = m.is_relaxation != 0;
    if (m.is_relaxation) {
        j["relaxation_step"] = m.relaxation_step;
        j["relaxation_total_steps"] = m.relaxation_total_steps;
        j["accumulated_time"] = m.accumulated_time;
        j["alpha_scaling"] = m.alpha_scaling;
        j["rho_center"] = m.rho_center;
        j["K"] = m.K;
        j["R"] = m.R;
        j["M_total"] = m.M_total;
    }
    j["kinetic_energy"] = m.kinetic_energy;
    j["thermal_energy"] = m.thermal_energy;
    j["potential_energy"] = m.potential_energy;
    j["total_energy"] = m.total_energy;
    return j;
}

void set_str(char *dst, size_t cap, const std::string &s)
{
    const size_t k = std::min(cap - 1, s.size());
    std::memcpy(dst, s.data(), k);
    dst[k] = '\0';
}

// OutputMetadata::from_json with OutputMetadata() / UnitSystem() defaults
void meta_from_json(const json &j, sphgpu_output_meta &m)
{
    static const char *const names[5][6] = {
        {"Code", "code_length", "code_mass", "code_time", "code_velocity", "code_energy"},
        {"Galactic", "kpc", "M_sun", "Myr", "km/s", "erg"},
        {"SI", "m", "kg", "s", "m/s", "J"},
        {"CGS", "cm", "g", "s", "cm/s", "erg"},
        {"Unknown", "unknown", "unknown", "unknown", "unknown", "unknown"}};
    std::memset(&m, 0, sizeof(m));
    set_str(m.simulation_name, sizeof(m.simulation_name), j.value("simulation_name", std::string("SPH Simulation")));
    // OutputMetadata::generate_timestamp: the current UTC time
    char now[32];
    const std::time_t t_now = std::time(nullptr);
    std::strftime(now, sizeof(now), "%Y-%m-%dT%H:%M:%SZ", std::gmtime(&t_now));
    set_str(m.timestamp, sizeof(m.timestamp), j.value("timestamp", std::string(now)));
    m.format_version = j.value("format_version", 1);
    m.step = j.value("step", 0);
    m.time_code = j.value("time_code", 0.0);
    m.time_physical = j.value("time_physical", 0.0);
    m.particle_count = j.value("particle_count", 0);
    m.sph_type = j.value("sph_type_enum", 0);
    m.kernel = j.value("kernel_type_enum", 0);
    m.gamma = j.value("gamma", 1.4);
    m.gravitational_constant = j.value("gravitational_constant", 1.0);
    m.neighbor_number = j.value("neighbor_number", 32);
    m.use_gravity = j.value("use_gravity", false);
    m.use_balsara = j.value("use_balsara", true);
    m.use_time_dependent_av = j.value("use_time_dependent_av", false);
    int t = 0;
    m.length_to_cgs = m.mass_to_cgs = m.time_to_cgs = m.velocity_to_cgs = 1.0;
    if (j.find("units") != j.end()) {
        const json &u = j["units"];
        t = u.value("type_enum", 0);
        m.length_to_cgs = u.value("length_to_cgs", 1.0);
        m.mass_to_cgs = u.value("mass_to_cgs", 1.0);
        m.time_to_cgs = u.value("time_to_cgs", 1.0);
        m.velocity_to_cgs = u.value("velocity_to_cgs", 1.0);
    }
    const char *const *nm = names[(t < 0 || t > 3) ? 4 : t];
    set_str(m.unit_type, sizeof(m.unit_type), nm[0]);
    set_str(m.length_unit, sizeof(m.length_unit), nm[1]);
    set_str(m.mass_unit, sizeof(m.mass_unit), nm[2]);
    set_str(m.time_unit, sizeof(m.time_unit), nm[3]);
    set_str(m.velocity_unit, sizeof(m.velocity_unit), nm[4]);
    set_str(m.energy_unit, sizeof(m.energy_unit), nm[5]);
    m.is_relaxation = j.value("is_relaxation", false);
    m.alpha_scaling = m.rho_center = m.K = m.R = m.M_total = 1.0;
    if (m.is_relaxation) {
        m.relaxation_step = j.value("relaxation_step", 0);
        m.relaxation_total_steps = j.value("relaxation_total_steps", 0);
        m.accumulated_time = j.value("accumulated_time", 0.0);
        m.alpha_scaling = j.value("alpha_scaling", 1.0);
        m.rho_center = j.value("rho_center", 0.0);
        m.K = j.value("K", 0.0);
        m.R = j.value("R", 0.0);
        m.M_total = j.value("M_total", 0.0);
    }
    m.kinetic_energy = j.value("kinetic_energy", 0.0);
    m.thermal_energy = j.value("thermal_energy", 0.0);
    m.potential_energy = j.value("potential_energy", 0.0);
    m.total_energy = j.value("total_energy", 0.0);
    m.precision = 16;
}

template <typename T> T swap_bytes(T v)
{
    unsigned char b[sizeof(T)], r[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    for (size_t i = 0; i < sizeof(T); ++i) r[i] = b[sizeof(T) - 1 - i];
    std::memcpy(&v, r, sizeof(T));
    return v;
}

bool str_attr(hid_t loc, const char *name, const std::string &v)
{
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, v.size() + 1);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    hid_t a = H5Acreate(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    if (a < 0) {
        H5Tclose(t);
        H5Sclose(sp);
        return false;
    }
    H5Awrite(a, t, v.c_str());
    H5Aclose(a);
    H5Tclose(t);
    H5Sclose(sp);
    return true;
}

// ---- describe: every object of a file as sorted text lines -----------------

std::string type_text(hid_t t)
{
    std::ostringstream o;
    const H5T_class_t c = H5Tget_class(t);
    o << "class=" << (int)c << " size=" << H5Tget_size(t);
    if (c == H5T_INTEGER) o << " sign=" << (int)H5Tget_sign(t) << " order=" << (int)H5Tget_order(t);
    if (c == H5T_FLOAT) o << " order=" << (int)H5Tget_order(t);
    if (c == H5T_STRING) o << " pad=" << (int)H5Tget_strpad(t) << " cset=" << (int)H5Tget_cset(t)
                           << " var=" << (int)H5Tis_variable_str(t);
    return o.str();
}

std::string space_text(hid_t sp)
{
    std::ostringstream o;
    const H5S_class_t c = H5Sget_simple_extent_type(sp);
    o << "space=" << (int)c;
    if (c == H5S_SIMPLE) {
        const int nd = H5Sget_simple_extent_ndims(sp);
        std::vector<hsize_t> d((size_t)nd), mx((size_t)nd);
        H5Sget_simple_extent_dims(sp, d.data(), mx.data());
        o << " dims=";
        for (int k = 0; k < nd; ++k) o << d[(size_t)k] << (k + 1 < nd ? "x" : "");
        o << " max=";
        for (int k = 0; k < nd; ++k) o << (long long)mx[(size_t)k] << (k + 1 < nd ? "x" : "");
    }
    return o.str();
}

// values of an attribute or dataset, in the native type of its class
std::string values_text(hid_t t, hid_t sp, bool is_attr, hid_t obj)
{
    std::ostringstream o;
    const H5T_class_t c = H5Tget_class(t);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    o.precision(17);
    if (c == H5T_STRING) {
        std::vector<char> buf(H5Tget_size(t) + 1, '\0');
        if (is_attr) H5Aread(obj, t, buf.data());
        else H5Dread(obj, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
        o << " value=" << std::string(buf.data());
    } else if (c == H5T_INTEGER) {
        std::vector<long long> v((size_t)n);
        if (is_attr) H5Aread(obj, H5T_NATIVE_LLONG, v.data());
        else H5Dread(obj, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        o << " values=";
        for (auto x : v) o << x << ",";
    } else if (c == H5T_FLOAT) {
        std::vector<double> v((size_t)n);
        if (is_attr) H5Aread(obj, H5T_NATIVE_DOUBLE, v.data());
        else H5Dread(obj, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        o << " values=";
        for (auto x : v) o << x << ",";
    }
    return o.str();
}

struct Walk {
    std::vector<std::string> lines;
};

herr_t attr_cb(hid_t loc, const char *name, const H5A_info_t *, void *op)
{
    Walk &w = *(Walk *)op;
    hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    hid_t t = H5Aget_type(a), sp = H5Aget_space(a);
    char path[512];
    H5Iget_name(loc, path, sizeof(path));
    w.lines.push_back(std::string("attr ") + path + "@" + name + " " + type_text(t) + " " + space_text(sp) +
                      values_text(t, sp, true, a));
    H5Sclose(sp);
    H5Tclose(t);
    H5Aclose(a);
    return 0;
}

herr_t obj_cb(hid_t root, const char *name, const H5O_info_t *info, void *op)
{
    Walk &w = *(Walk *)op;
    const std::string path = std::string("/") + (std::strcmp(name, ".") ? name : "");
    if (info->type == H5O_TYPE_GROUP) {
        w.lines.push_back("group " + path);
        hid_t g = H5Gopen(root, name, H5P_DEFAULT);
        H5Aiterate2(g, H5_INDEX_NAME, H5_ITER_INC, nullptr, attr_cb, op);
        H5Gclose(g);
    } else if (info->type == H5O_TYPE_DATASET) {
        hid_t d = H5Dopen(root, name, H5P_DEFAULT);
        hid_t t = H5Dget_type(d), sp = H5Dget_space(d), pl = H5Dget_create_plist(d);
        std::ostringstream o;
        o << "dataset " << path << " " << type_text(t) << " " << space_text(sp) << " layout=" << (int)H5Pget_layout(pl);
        if (H5Pget_layout(pl) == H5D_CHUNKED) {
            hsize_t ch[4] = {0, 0, 0, 0};
            const int nd = H5Pget_chunk(pl, 4, ch);
            o << " chunk=";
            for (int k = 0; k < nd; ++k) o << ch[k] << (k + 1 < nd ? "x" : "");
        }
        const int nf = H5Pget_nfilters(pl);
        for (int k = 0; k < nf; ++k) {
            unsigned flags = 0, cd[8];
            size_t ncd = 8;
            char fname[64];
            const H5Z_filter_t id = H5Pget_filter2(pl, (unsigned)k, &flags, &ncd, cd, sizeof(fname), fname, nullptr);
            o << " filter=" << (int)id << "(";
            for (size_t q = 0; q < ncd; ++q) o << cd[q] << ",";
            o << ")";
        }
        o << values_text(t, sp, false, d);
        w.lines.push_back(o.str());
        H5Aiterate2(d, H5_INDEX_NAME, H5_ITER_INC, nullptr, attr_cb, op);
        H5Pclose(pl);
        H5Sclose(sp);
        H5Tclose(t);
        H5Dclose(d);
    }
    return 0;
}

} // namespace

extern "C" {

// to_json().dump(2) of the metadata (which = 0) or of its unit system (1)
int orc_meta_json(const sphgpu_output_meta *m, int which, char *buf, int len)
{
    const std::string s = (which == 0 ? meta_json(*m) : units_json(*m)).dump(2);
    if ((int)s.size() + 1 > len) return -1;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

int orc_meta_from_json(const char *text, sphgpu_output_meta *m)
{
    try {
        meta_from_json(json::parse(text), *m);
    } catch (...) {
        return -1;
    }
    return 0;
}

// VTKWriter::open + write_particles; -1 where write_particles returns false
int orc_vtk_write(const char *path, const sphgpu_output_meta *meta, int dim, int n, const double *f, int binary)
{
    std::ofstream out(path, binary ? (std::ios::out | std::ios::binary) : std::ios::out);
    if (!out.is_open()) return -1;
    const bool bin = binary != 0;
    out << "# vtk DataFile Version 3.0\n";
    out << "SPH Simulation - " << meta->simulation_name << "\n";
    out << (bin ? "BINARY\n" : "ASCII\n");
    out << "DATASET UNSTRUCTURED_GRID\n";
    if (n == 0) return -1;
    auto F = [&](int row, int i) { return f[(size_t)row * (size_t)n + (size_t)i]; };
    auto wf = [&](double v) {
        const float x = swap_bytes((float)v);
        out.write((const char *)&x, sizeof(float));
    };
    auto wi = [&](int v) {
        const int x = swap_bytes(v);
        out.write((const char *)&x, sizeof(int));
    };
    auto vec = [&](int base) {
        for (int i = 0; i < n; ++i) {
            if (bin) {
                wf(F(base, i));
                wf(dim >= 2 ? F(base + 1, i) : 0.0);
                wf(dim >= 3 ? F(base + 2, i) : 0.0);
            } else {
                out << F(base, i);
                if (dim >= 2) out << " " << F(base + 1, i);
                else out << " 0.0";
                if (dim >= 3) out << " " << F(base + 2, i);
                else out << " 0.0";
                out << "\n";
            }
        }
        if (!bin) out << "\n";
    };
    out << "POINTS " << n << " float\n";
    vec(SPHGPU_F_POS);
    out << "CELLS " << n << " " << (n * 2) << "\n";
    for (int i = 0; i < n; ++i) {
        if (bin) {
            wi(1);
            wi(i);
        } else {
            out << "1 " << i << "\n";
        }
    }
    out << "CELL_TYPES " << n << "\n";
    for (int i = 0; i < n; ++i) {
        if (bin) wi(1);
        else out << "1\n";
    }
    out << "POINT_DATA " << n << "\n";
    const std::pair<const char *, int> sc[] = {{"density", SPHGPU_F_DENS},
                                               {"pressure", SPHGPU_F_PRES},
                                               {"mass", SPHGPU_F_MASS},
                                               {"internal_energy", SPHGPU_F_ENE},
                                               {"smoothing_length", SPHGPU_F_SML},
                                               {"sound_speed", SPHGPU_F_SOUND}};
    for (const auto &s : sc) {
        out << "SCALARS " << s.first << " float 1\n";
        out << "LOOKUP_TABLE default\n";
        for (int i = 0; i < n; ++i) {
            if (bin) wf(F(s.second, i));
            else out << F(s.second, i) << "\n";
        }
        if (!bin) out << "\n";
    }
    out << "SCALARS particle_id int 1\n";
    out << "LOOKUP_TABLE default\n";
    for (int i = 0; i < n; ++i) {
        const int id = (int)F(SPHGPU_F_ID, i);
        if (bin) wi(id);
        else out << id << "\n";
    }
    if (!bin) out << "\n";
    out << "VECTORS velocity float\n";
    vec(SPHGPU_F_VEL);
    out << "VECTORS acceleration float\n";
    vec(SPHGPU_F_ACC);
    return out.good() ? 0 : -1;
}

// HDF5Writer::open + write_particles + close
int orc_hdf5_write(const char *path, const sphgpu_output_meta *m, int dim, int n, const double *f, int compression)
{
    int comp = compression < 0 ? 0 : (compression > 9 ? 9 : compression);
    hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return -1;
    // write_metadata_group
    hid_t g = H5Gcreate(file, "/metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t one[1] = {1};
    hid_t sp = H5Screate_simple(1, one, nullptr);
    auto sa = [&](const char *name, hid_t t, const void *v) {
        hid_t a = H5Acreate(g, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
        H5Awrite(a, t, v);
        H5Aclose(a);
    };
    sa("format_version", H5T_NATIVE_INT, &m->format_version);
    sa("step", H5T_NATIVE_INT, &m->step);
    sa("particle_count", H5T_NATIVE_INT, &m->particle_count);
    sa("time_code", H5T_NATIVE_DOUBLE, &m->time_code);
    sa("time_physical", H5T_NATIVE_DOUBLE, &m->time_physical);
    sa("gamma", H5T_NATIVE_DOUBLE, &m->gamma);
    H5Sclose(sp);
    str_attr(g, "simulation_name", m->simulation_name);
    str_attr(g, "timestamp", m->timestamp);
    str_attr(g, "sph_type", type_name(m->sph_type));
    str_attr(g, "kernel_type", kern_name(m->kernel));
    str_attr(g, "unit_system", units_json(*m).dump(2));
    str_attr(g, "metadata_json", meta_json(*m).dump(2));
    H5Gclose(g);
    // write_energy_group
    g = H5Gcreate(file, "/energy", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    sp = H5Screate_simple(1, one, nullptr);
    const std::pair<const char *, const double *> e[] = {{"kinetic", &m->kinetic_energy},
                                                         {"thermal", &m->thermal_energy},
                                                         {"potential", &m->potential_energy},
                                                         {"total", &m->total_energy}};
    for (const auto &x : e) {
        hid_t d = H5Dcreate(g, x.first, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, x.second);
        H5Dclose(d);
    }
    H5Sclose(sp);
    H5Gclose(g);
    if (n == 0) {
        H5Fclose(file);
        return -1;
    }
    // write_particles_group
    const hsize_t N = (hsize_t)n;
    g = H5Gcreate(file, "/particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hid_t plist = H5P_DEFAULT;
    if (comp > 0) {
        plist = H5Pcreate(H5P_DATASET_CREATE);
        hsize_t ch[1] = {N < 1024 ? N : 1024};
        H5Pset_chunk(plist, 1, ch);
        H5Pset_deflate(plist, (unsigned)comp);
    }
    auto wd = [&](const char *name, const std::vector<double> &v) {
        hsize_t d1[1] = {N};
        hid_t s = H5Screate_simple(1, d1, nullptr);
        hid_t d = H5Dcreate(g, name, H5T_NATIVE_DOUBLE, s, H5P_DEFAULT, plist, H5P_DEFAULT);
        H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        H5Dclose(d);
        H5Sclose(s);
    };
    auto wi = [&](const char *name, const std::vector<int> &v) {
        hsize_t d1[1] = {N};
        hid_t s = H5Screate_simple(1, d1, nullptr);
        hid_t pi = H5P_DEFAULT;
        if (comp > 0) {
            pi = H5Pcreate(H5P_DATASET_CREATE);
            hsize_t ch[1] = {N < 1024 ? N : 1024};
            H5Pset_chunk(pi, 1, ch);
            H5Pset_deflate(pi, (unsigned)comp);
        }
        hid_t d = H5Dcreate(g, name, H5T_NATIVE_INT, s, H5P_DEFAULT, pi, H5P_DEFAULT);
        H5Dwrite(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        H5Dclose(d);
        if (pi != H5P_DEFAULT) H5Pclose(pi);
        H5Sclose(s);
    };
    auto col = [&](int row) {
        std::vector<double> v((size_t)n, 0.0);
        if (row >= 0)
            for (int i = 0; i < n; ++i) v[(size_t)i] = f[(size_t)row * (size_t)n + (size_t)i];
        return v;
    };
    auto icol = [&](int row) {
        std::vector<int> v((size_t)n);
        for (int i = 0; i < n; ++i) v[(size_t)i] = (int)f[(size_t)row * (size_t)n + (size_t)i];
        return v;
    };
    wi("id", icol(SPHGPU_F_ID));
    const char *vn[9] = {"pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x", "acc_y", "acc_z"};
    const int vb[3] = {SPHGPU_F_POS, SPHGPU_F_VEL, SPHGPU_F_ACC};
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) wd(vn[3 * v + a], col(a < dim ? vb[v] + a : -1));
    wd("mass", col(SPHGPU_F_MASS));
    wd("dens", col(SPHGPU_F_DENS));
    wd("pres", col(SPHGPU_F_PRES));
    wd("ene", col(SPHGPU_F_ENE));
    wd("sml", col(SPHGPU_F_SML));
    wd("sound", col(SPHGPU_F_SOUND));
    wd("alpha", col(SPHGPU_F_ALPHA));
    wd("balsara", col(SPHGPU_F_BALSARA));
    wd("gradh", col(SPHGPU_F_GRADH));
    wd("phi", col(SPHGPU_F_PHI));
    wi("neighbor", icol(SPHGPU_F_NEIGHBOR));
    if (plist != H5P_DEFAULT) H5Pclose(plist);
    H5Gclose(g);
    H5Fclose(file);
    return 0;
}

// HDF5Writer::read_particles + read_metadata: particles in file order, placed
// by id into the block (n_cap rows); *rows = the file's count
int orc_hdf5_read(const char *path, sphgpu_output_meta *m, int n_cap, double *f, int *rows)
{
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return -1;
    hid_t g = H5Gopen(file, "/particles", H5P_DEFAULT);
    if (g < 0) {
        H5Fclose(file);
        return -1;
    }
    hid_t d = H5Dopen(g, "id", H5P_DEFAULT);
    if (d < 0) {
        H5Gclose(g);
        H5Fclose(file);
        return -1;
    }
    hid_t s = H5Dget_space(d);
    hsize_t dims[1];
    H5Sget_simple_extent_dims(s, dims, nullptr);
    const hsize_t N = dims[0];
    H5Sclose(s);
    H5Dclose(d);
    *rows = (int)N;
    auto rd = [&](const char *name, std::vector<double> &v) {
        v.assign(N, 0.0);
        hid_t x = H5Dopen(g, name, H5P_DEFAULT);
        if (x < 0) return;
        H5Dread(x, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        H5Dclose(x);
    };
    auto ri = [&](const char *name, std::vector<int> &v) {
        v.assign(N, 0);
        hid_t x = H5Dopen(g, name, H5P_DEFAULT);
        if (x < 0) return;
        H5Dread(x, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        H5Dclose(x);
    };
    std::vector<int> ids, nb;
    ri("id", ids);
    ri("neighbor", nb);
    const char *names[19] = {"pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x",
                             "acc_y", "acc_z", "mass",  "dens",  "pres",  "ene",   "sml",
                             "sound", "alpha", "balsara", "gradh", "phi"};
    const int rowsof[19] = {SPHGPU_F_POS,   SPHGPU_F_POS + 1, SPHGPU_F_POS + 2, SPHGPU_F_VEL,   SPHGPU_F_VEL + 1,
                            SPHGPU_F_VEL + 2, SPHGPU_F_ACC, SPHGPU_F_ACC + 1, SPHGPU_F_ACC + 2, SPHGPU_F_MASS,
                            SPHGPU_F_DENS,  SPHGPU_F_PRES,  SPHGPU_F_ENE,    SPHGPU_F_SML,   SPHGPU_F_SOUND,
                            SPHGPU_F_ALPHA, SPHGPU_F_BALSARA, SPHGPU_F_GRADH, SPHGPU_F_PHI};
    for (hsize_t i = 0; i < N; ++i) {
        if (ids[i] < 0 || ids[i] >= n_cap) {
            H5Gclose(g);
            H5Fclose(file);
            return -1;
        }
    }
    std::vector<double> v;
    for (int k = 0; k < 19; ++k) {
        rd(names[k], v);
        for (hsize_t i = 0; i < N; ++i) f[(size_t)rowsof[k] * (size_t)n_cap + (size_t)ids[i]] = v[i];
    }
    for (hsize_t i = 0; i < N; ++i) {
        f[(size_t)SPHGPU_F_ID * (size_t)n_cap + (size_t)ids[i]] = ids[i];
        f[(size_t)SPHGPU_F_NEIGHBOR * (size_t)n_cap + (size_t)ids[i]] = nb[i];
    }
    H5Gclose(g);
    // read_metadata
    g = H5Gopen(file, "/metadata", H5P_DEFAULT);
    int rc = 0;
    if (g < 0) {
        rc = -1;
    } else {
        if (H5Aexists(g, "metadata_json") > 0) {
            hid_t a = H5Aopen(g, "metadata_json", H5P_DEFAULT);
            hid_t t = H5Aget_type(a);
            std::vector<char> buf(H5Tget_size(t) + 1, '\0');
            H5Aread(a, t, buf.data());
            H5Tclose(t);
            H5Aclose(a);
            try {
                meta_from_json(json::parse(std::string(buf.data())), *m);
            } catch (...) {
                rc = -1;
            }
        } else {
            meta_from_json(json::object(), *m);
        }
        H5Gclose(g);
    }
    H5Fclose(file);
    return rc;
}

// every group, attribute and dataset of a file, one sorted line each
int orc_hdf5_describe(const char *path, char *buf, int len)
{
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return -1;
    Walk w;
    H5Ovisit(file, H5_INDEX_NAME, H5_ITER_INC, obj_cb, &w);
    H5Fclose(file);
    std::sort(w.lines.begin(), w.lines.end());
    std::string s;
    for (const auto &l : w.lines) s += l + "\n";
    if ((int)s.size() + 1 > len) return -(int)(s.size() + 1);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

} // extern "C"
