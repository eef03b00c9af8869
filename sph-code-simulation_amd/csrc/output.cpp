// output.cpp -- snapshot CSV files and the energy log in the reference's
// on-disk layout.
//
// Reference: CSVWriter (src/writers/csv_writer.cpp:17-165 writer, 174-371
// reader), OutputManager (src/output_manager.cpp:55-73 energy header, 235-252
// energy rows, 254-327 metadata and energies, 329-338 file names) and
// OutputMetadata (src/output_metadata.cpp:8-65).  The reference's streams are
// set to std::scientific with setprecision(p), which prints a double as the C
// conversion "%.{p}e" does; integers print in decimal.  The particle rows are
// formatted by host threads into per-thread buffers (rows in id order) and
// written with one call per buffer; the reader splits the file at line
// boundaries and parses the chunks in parallel.
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sphgpu.h"

namespace {

constexpr int kColumns = 21;
constexpr int kHeaderLines = 50;  // the reader skips 50 lines + the column line

int worker_count(long long items)
{
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    const long long by_size = std::max(1ll, items / 4096);
    return (int)std::min<long long>(std::min<unsigned>(hw, 16u), by_size);
}

template <typename F> void parallel_ranges(long long n, F f)
{
    const int t = worker_count(n);
    if (t <= 1) {
        f(0, 0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int k = 0; k < t; ++k) {
        const long long a = n * k / t, b = n * (k + 1) / t;
        th.emplace_back([=, &f] { f(k, a, b); });
    }
    for (auto &x : th) x.join();
}

void put_double(std::string &s, double v, int prec)
{
    char buf[64];
    const int k = std::snprintf(buf, sizeof(buf), "%.*e", prec, v);
    s.append(buf, (size_t)k);
}

void put_int(std::string &s, long long v)
{
    char buf[32];
    const int k = std::snprintf(buf, sizeof(buf), "%lld", v);
    s.append(buf, (size_t)k);
}

const char *sph_type_name(int t)
{
    switch (t) {
    case SPHGPU_SSPH: return "Standard SPH";
    case SPHGPU_DISPH: return "Density Independent SPH";
    case SPHGPU_GSPH: return "Godunov SPH";
    default: return "Unknown";
    }
}

const char *kernel_name(int k)
{
    switch (k) {
    case SPHGPU_CUBIC_SPLINE: return "Cubic Spline";
    case SPHGPU_WENDLAND: return "Wendland C4";
    default: return "Unknown";
    }
}

void copy_str(char *dst, size_t cap, const std::string &src)
{
    const size_t k = std::min(cap - 1, src.size());
    std::memcpy(dst, src.data(), k);
    dst[k] = '\0';
}

std::string utc_timestamp()
{
    const std::time_t now = std::time(nullptr);
    std::tm t;
    gmtime_r(&now, &t);
    char buf[32];
    std::strftime(buf, sizeof(buf), "%Y-%m-%dT%H:%M:%SZ", &t);
    return buf;
}

// OutputMetadata() defaults (src/output_metadata.cpp:8-37) and UnitSystem()
// code units (src/units.cpp:6-13, 125-205)
void meta_defaults(sphgpu_output_meta *m)
{
    std::memset(m, 0, sizeof(*m));
    m->format_version = 1;
    copy_str(m->timestamp, sizeof(m->timestamp), utc_timestamp());
    copy_str(m->simulation_name, sizeof(m->simulation_name), "SPH Simulation");
    copy_str(m->unit_type, sizeof(m->unit_type), "Code");
    copy_str(m->length_unit, sizeof(m->length_unit), "code_length");
    copy_str(m->mass_unit, sizeof(m->mass_unit), "code_mass");
    copy_str(m->time_unit, sizeof(m->time_unit), "code_time");
    copy_str(m->velocity_unit, sizeof(m->velocity_unit), "code_velocity");
    copy_str(m->energy_unit, sizeof(m->energy_unit), "code_energy");
    m->length_to_cgs = m->mass_to_cgs = m->time_to_cgs = m->velocity_to_cgs = 1.0;
    m->sph_type = SPHGPU_SSPH;
    m->kernel = SPHGPU_CUBIC_SPLINE;
    m->gamma = 1.4;
    m->gravitational_constant = 1.0;
    m->neighbor_number = 32;
    m->use_balsara = 1;
    m->alpha_scaling = 1.0;
    m->rho_center = 1.0;
    m->K = 1.0;
    m->R = 1.0;
    m->M_total = 1.0;
    m->precision = 16;
}

// CSVWriter::write_header + write_column_names (csv_writer.cpp:42-130)
std::string csv_header(const sphgpu_output_meta &m)
{
    const int p = m.precision;
    std::string s;
    auto line = [&](const std::string &t) { s += t; s += '\n'; };
    auto dbl = [&](double v) { std::string t; put_double(t, v, p); return t; };
    auto num = [&](long long v) { std::string t; put_int(t, v); return t; };
    line("# SPH Simulation Output - CSV Format v" + num(m.format_version));
    line(std::string("# Timestamp: ") + m.timestamp);
    line(std::string("# Simulation: ") + m.simulation_name);
    line("#");
    line("# === Unit System ===");
    line(std::string("# Type: ") + m.unit_type);
    line(std::string("# Length: ") + m.length_unit + " (" + dbl(m.length_to_cgs) + " cm)");
    line(std::string("# Mass: ") + m.mass_unit + " (" + dbl(m.mass_to_cgs) + " g)");
    line(std::string("# Time: ") + m.time_unit + " (" + dbl(m.time_to_cgs) + " s)");
    line(std::string("# Velocity: ") + m.velocity_unit + " (" + dbl(m.velocity_to_cgs) + " cm/s)");
    line("#");
    line("# === Simulation State ===");
    line("# Step: " + num(m.step));
    line("# Time (code): " + dbl(m.time_code));
    line("# Time (physical): " + dbl(m.time_physical) + " " + m.time_unit);
    line("# Particle Count: " + num(m.particle_count));
    line("#");
    line("# === Physics Parameters ===");
    line("# Gamma: " + dbl(m.gamma));
    line("# G: " + dbl(m.gravitational_constant));
    line("# Neighbor Number: " + num(m.neighbor_number));
    line(std::string("# SPH Type: ") + sph_type_name(m.sph_type));
    line(std::string("# Kernel: ") + kernel_name(m.kernel));
    line(std::string("# Gravity: ") + (m.use_gravity ? "enabled" : "disabled"));
    line(std::string("# Balsara Switch: ") + (m.use_balsara ? "enabled" : "disabled"));
    line(std::string("# Time-Dependent AV: ") + (m.use_time_dependent_av ? "enabled" : "disabled"));
    line("#");
    line("# === Energy ===");
    line("# Kinetic: " + dbl(m.kinetic_energy) + " [code units]");
    line("# Thermal: " + dbl(m.thermal_energy) + " [code units]");
    line("# Potential: " + dbl(m.potential_energy) + " [code units]");
    line("# Total: " + dbl(m.total_energy) + " [code units]");
    line("#");
    if (m.is_relaxation) {
        line("# === Lane-Emden Relaxation ===");
        line("# Relaxation Step: " + num(m.relaxation_step) + " / " + num(m.relaxation_total_steps));
        line("# Accumulated Time: " + dbl(m.accumulated_time));
        line("# Alpha Scaling: " + dbl(m.alpha_scaling));
        line("# Central Density: " + dbl(m.rho_center));
        line("# Polytropic K: " + dbl(m.K));
        line("# Radius: " + dbl(m.R));
        line("# Total Mass: " + dbl(m.M_total));
        line("#");
    }
    line("# === Columns ===");
    line("# id: Particle ID");
    line("# pos_x, pos_y, pos_z: Position [code units]");
    line("# vel_x, vel_y, vel_z: Velocity [code units]");
    line("# acc_x, acc_y, acc_z: Acceleration [code units]");
    line("# mass: Particle mass [code units]");
    line("# dens: Density [code units]");
    line("# pres: Pressure [code units]");
    line("# ene: Specific internal energy [code units]");
    line("# sml: Smoothing length [code units]");
    line("# sound: Sound speed [code units]");
    line("# alpha: Artificial viscosity coefficient");
    line("# balsara: Balsara factor");
    line("# gradh: grad-h term factor");
    line("# phi: Gravitational potential [code units]");
    line("# neighbor: Number of neighbors");
    line("#");
    line("id,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z,acc_x,acc_y,acc_z,mass,dens,pres,ene,sml,sound,"
         "alpha,balsara,gradh,phi,neighbor");
    return s;
}

// the column order of a row (csv_writer.cpp:139-160) as field rows of the
// [SPHGPU_NFIELDS][n] block; -1 = an axis the context does not have
void row_fields(int dim, int f[kColumns])
{
    int k = 0;
    f[k++] = SPHGPU_F_ID;
    for (int base : {SPHGPU_F_POS, SPHGPU_F_VEL, SPHGPU_F_ACC})
        for (int a = 0; a < 3; ++a) f[k++] = a < dim ? base + a : -1;
    for (int g : {SPHGPU_F_MASS, SPHGPU_F_DENS, SPHGPU_F_PRES, SPHGPU_F_ENE, SPHGPU_F_SML, SPHGPU_F_SOUND,
                  SPHGPU_F_ALPHA, SPHGPU_F_BALSARA, SPHGPU_F_GRADH, SPHGPU_F_PHI, SPHGPU_F_NEIGHBOR})
        f[k++] = g;
}

bool parse_int(const std::string &s, int &out)
{
    // std::stoi: leading number, errors on no digits or out of int range
    errno = 0;
    char *end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || errno == ERANGE || v < INT_MIN || v > INT_MAX) return false;
    out = (int)v;
    return true;
}

bool parse_double(const char *b, const char *e, double &out)
{
    // std::stod: leading number, errors on no conversion or ERANGE
    char tmp[128];
    const size_t k = std::min((size_t)(e - b), sizeof(tmp) - 1);
    std::memcpy(tmp, b, k);
    tmp[k] = '\0';
    errno = 0;
    char *end = nullptr;
    out = std::strtod(tmp, &end);
    return end != tmp && errno != ERANGE;
}

bool parse_double(const std::string &s, double &out) { return parse_double(s.data(), s.data() + s.size(), out); }

std::string trim(const std::string &x)
{
    const size_t a = x.find_first_not_of(" \t");
    if (a == std::string::npos) return "";
    const size_t b = x.find_last_not_of(" \t");
    return x.substr(a, b - a + 1);
}

// CSVWriter::read_metadata (csv_writer.cpp:174-269); false where a value of a
// known key does not parse (std::stoi / std::stod would throw there)
bool read_header_fields(const std::vector<std::string> &lines, sphgpu_output_meta &m)
{
    bool ok = true;
    for (size_t ln = 0; ln < lines.size() && ln < (size_t)kHeaderLines; ++ln) {
        const std::string &line = lines[ln];
        if (line.empty() || line[0] != '#') continue;
        const size_t colon = line.find(':');
        if (colon == std::string::npos || colon < 2) continue;
        const std::string key = trim(line.substr(2, colon - 2));
        const std::string value = trim(colon + 2 <= line.size() ? line.substr(colon + 2) : std::string());
        int iv;
        double dv;
        auto geti = [&](int32_t &dst) { if (parse_int(value, iv)) dst = iv; else ok = false; };
        auto getd = [&](double &dst) { if (parse_double(value, dv)) dst = dv; else ok = false; };
        auto before_bracket = [&](double &dst) {
            const size_t b = value.find(" [");
            if (b == std::string::npos) return;
            if (parse_double(value.substr(0, b), dv)) dst = dv;
            else ok = false;
        };
        if (key == "Timestamp") copy_str(m.timestamp, sizeof(m.timestamp), value);
        else if (key == "Simulation") copy_str(m.simulation_name, sizeof(m.simulation_name), value);
        else if (key == "Step") geti(m.step);
        else if (key == "Time (code)") getd(m.time_code);
        else if (key == "Particle Count") geti(m.particle_count);
        else if (key == "Gamma") getd(m.gamma);
        else if (key == "G") getd(m.gravitational_constant);
        else if (key == "Neighbor Number") geti(m.neighbor_number);
        else if (key == "Kinetic") before_bracket(m.kinetic_energy);
        else if (key == "Thermal") before_bracket(m.thermal_energy);
        else if (key == "Potential") before_bracket(m.potential_energy);
        else if (key == "Total") before_bracket(m.total_energy);
        else if (key == "Relaxation Step") {
            m.is_relaxation = 1;
            const size_t sl = value.find('/');
            if (sl != std::string::npos) {
                if (parse_int(value.substr(0, sl), iv)) m.relaxation_step = iv;
                else ok = false;
                if (sl + 2 <= value.size() && parse_int(value.substr(sl + 2), iv)) m.relaxation_total_steps = iv;
                else ok = false;
            }
        }
        else if (key == "Accumulated Time") getd(m.accumulated_time);
        else if (key == "Alpha Scaling") getd(m.alpha_scaling);
        else if (key == "Central Density") getd(m.rho_center);
        else if (key == "Polytropic K") getd(m.K);
        else if (key == "Radius") getd(m.R);
        else if (key == "Total Mass") getd(m.M_total);
    }
    return ok;
}

} // namespace

extern "C" {

int sphgpu_output_meta_init(sphgpu_ctx *ctx, int32_t step, sphgpu_output_meta *m)
{
    if (!ctx || !m) return SPHGPU_E_ARG;
    meta_defaults(m);
    int rc;
    int32_t n = 0;
    sphgpu_params p;
    if ((rc = sphgpu_get_params(ctx, &p, &n))) return rc;
    double t = 0.0, e[7];
    if ((rc = sphgpu_get_time(ctx, &t))) return rc;
    if ((rc = sphgpu_energy_momentum(ctx, e))) return rc;
    m->step = step;
    m->time_code = t;
    m->time_physical = t * m->time_to_cgs;
    m->particle_count = n;
    m->sph_type = p.sph_type;
    m->kernel = p.kernel;
    m->gamma = p.gamma;
    m->gravitational_constant = p.gravity_constant;
    m->neighbor_number = p.neighbor_number;
    m->use_gravity = p.use_gravity;
    m->use_balsara = p.use_balsara_switch;
    m->use_time_dependent_av = p.use_time_dependent_av;
    m->kinetic_energy = e[0];
    m->thermal_energy = e[1];
    m->potential_energy = e[2];
    m->total_energy = e[0] + e[1] + e[2];
    return 0;
}

int sphgpu_snapshot_filename(const char *dir, const char *prefix, int32_t count, const char *ext, char *buf,
                             int32_t buf_len)
{
    if (!dir || !prefix || !ext || !buf || buf_len <= 0) return SPHGPU_E_ARG;
    const int k = std::snprintf(buf, (size_t)buf_len, "%s/%s_%04d%s", dir, prefix, count, ext);
    return (k < 0 || k >= buf_len) ? SPHGPU_E_ARG : 0;
}

int sphgpu_write_csv_fields(const char *path, const sphgpu_output_meta *meta, int32_t dim, int32_t n,
                            const double *f)
{
    if (!path || !meta || !f || n < 0 || dim < 1 || dim > 3 || meta->precision < 0 || meta->precision > 40)
        return SPHGPU_E_ARG;
    int col[kColumns];
    row_fields(dim, col);
    const int prec = meta->precision;
    const int t = worker_count(n);
    std::vector<std::string> part((size_t)std::max(t, 1));
    parallel_ranges(n, [&](int k, long long a, long long b) {
        std::string &s = part[(size_t)k];
        s.reserve((size_t)(b - a) * (size_t)(19 * (prec + 8) + 24));
        for (long long i = a; i < b; ++i) {
            for (int c = 0; c < kColumns; ++c) {
                const int fr = col[c];
                const double v = fr < 0 ? 0.0 : f[(size_t)fr * n + (size_t)i];
                if (c == 0 || c == kColumns - 1) put_int(s, (long long)v);
                else put_double(s, v, prec);
                s += (c == kColumns - 1) ? '\n' : ',';
            }
        }
    });
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return SPHGPU_E_ARG;
    const std::string head = csv_header(*meta);
    bool ok = std::fwrite(head.data(), 1, head.size(), fp) == head.size();
    for (const auto &s : part) ok = ok && std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : SPHGPU_E_ARG;
}

int sphgpu_write_snapshot_csv(sphgpu_ctx *ctx, const char *path, const sphgpu_output_meta *meta)
{
    if (!ctx || !path || !meta) return SPHGPU_E_ARG;
    int32_t n = 0;
    sphgpu_params p;
    int rc;
    if ((rc = sphgpu_get_params(ctx, &p, &n))) return rc;
    std::vector<double> f((size_t)SPHGPU_NFIELDS * n);
    if ((rc = sphgpu_download(ctx, f.data(), SPHGPU_MASK_ALL))) return rc;
    return sphgpu_write_csv_fields(path, meta, p.dim, n, f.data());
}

namespace {

bool slurp(const char *path, std::string &text)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return false;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, k);
    std::fclose(fp);
    return true;
}

// the 51 header lines and the start of every non-empty row after them
bool split_csv(const std::string &text, std::vector<std::string> &head, std::vector<size_t> &starts)
{
    size_t pos = 0;
    for (int ln = 0; ln < kHeaderLines + 1; ++ln) {
        if (pos >= text.size()) return false;  // fewer than 51 lines
        size_t e = text.find('\n', pos);
        if (e == std::string::npos) e = text.size();
        head.emplace_back(text.substr(pos, e - pos));
        pos = e + 1;
    }
    for (size_t q = pos; q < text.size();) {
        size_t e = text.find('\n', q);
        if (e == std::string::npos) e = text.size();
        if (e > q) starts.push_back(q);  // empty lines are skipped
        q = e + 1;
    }
    return true;
}

} // namespace

int sphgpu_csv_row_count(const char *path, int32_t *rows)
{
    if (!path || !rows) return SPHGPU_E_ARG;
    std::string text;
    std::vector<std::string> head;
    std::vector<size_t> starts;
    if (!slurp(path, text) || !split_csv(text, head, starts)) return SPHGPU_E_ARG;
    *rows = (int32_t)starts.size();
    return 0;
}

int sphgpu_read_csv_fields(const char *path, sphgpu_output_meta *meta, int32_t dim, int32_t n, double *f)
{
    if (!path || !f || n <= 0 || dim < 1 || dim > 3) return SPHGPU_E_ARG;
    std::string text;
    std::vector<std::string> head;
    std::vector<size_t> starts;
    if (!slurp(path, text) || !split_csv(text, head, starts)) return SPHGPU_E_ARG;
    sphgpu_output_meta m;
    meta_defaults(&m);
    if (!read_header_fields(head, m)) return SPHGPU_E_ARG;
    const long long rows = (long long)starts.size();
    if (rows == 0 || rows != n) return SPHGPU_E_ARG;
    int col[kColumns];
    row_fields(3, col);  // every column is read; axes beyond dim are dropped
    std::vector<double> vals((size_t)rows * kColumns);
    std::vector<char> bad((size_t)std::max(worker_count(rows), 1), 0);
    parallel_ranges(rows, [&](int k, long long a, long long b) {
        for (long long r = a; r < b && !bad[(size_t)k]; ++r) {
            const char *s = text.data() + starts[(size_t)r];
            const char *end = text.data() + text.size();
            const char *nl = static_cast<const char *>(std::memchr(s, '\n', (size_t)(end - s)));
            if (!nl) nl = end;
            // split on ',' as std::getline does: a trailing comma ends the
            // last field, so "a,b," has two fields
            int c = 0;
            const char *fb = s;
            for (const char *q = s;; ++q) {
                if (q == nl || *q == ',') {
                    if (!(q == nl && q == fb && c > 0 && *(q - 1) == ',')) {
                        if (c >= kColumns) { bad[(size_t)k] = 1; break; }
                        double v;
                        bool ok;
                        if (c == 0 || c == kColumns - 1) {
                            int iv = 0;
                            ok = parse_int(std::string(fb, q), iv);
                            v = iv;
                        } else {
                            ok = parse_double(fb, q, v);
                        }
                        if (!ok) { bad[(size_t)k] = 1; break; }
                        vals[(size_t)r * kColumns + c] = v;
                        ++c;
                    }
                    if (q == nl) break;
                    fb = q + 1;
                }
            }
            if (c != kColumns) bad[(size_t)k] = 1;
        }
    });
    for (char b : bad)
        if (b) return SPHGPU_E_ARG;
    // place by id (the reference indexes particles by id)
    std::fill(f, f + (size_t)SPHGPU_NFIELDS * n, 0.0);
    std::vector<char> seen((size_t)n, 0);
    for (long long r = 0; r < rows; ++r) {
        const double idv = vals[(size_t)r * kColumns];
        if (!(idv >= 0.0 && idv < (double)n)) return SPHGPU_E_ARG;
        const int id = (int)idv;
        if (seen[(size_t)id]) return SPHGPU_E_ARG;
        seen[(size_t)id] = 1;
        for (int c = 0; c < kColumns; ++c) {
            const int fr = col[c];
            const bool axis = fr < SPHGPU_F_MASS;  // pos, vel, acc rows
            if (axis && (fr % 3) >= dim) continue;
            f[(size_t)fr * n + (size_t)id] = vals[(size_t)r * kColumns + c];
        }
    }
    if (meta) *meta = m;
    return 0;
}

int sphgpu_read_snapshot_csv(sphgpu_ctx *ctx, const char *path, sphgpu_output_meta *meta)
{
    if (!ctx || !path) return SPHGPU_E_ARG;
    int32_t n = 0;
    sphgpu_params p;
    int rc;
    if ((rc = sphgpu_get_params(ctx, &p, &n))) return rc;
    std::vector<double> f((size_t)SPHGPU_NFIELDS * n);
    sphgpu_output_meta m;
    if ((rc = sphgpu_read_csv_fields(path, &m, p.dim, n, f.data()))) return rc;
    if ((rc = sphgpu_upload(ctx, f.data(), SPHGPU_MASK_ALL))) return rc;
    if ((rc = sphgpu_set_time(ctx, m.time_code))) return rc;
    if (meta) *meta = m;
    return 0;
}

int sphgpu_energy_file_begin(const char *path, const sphgpu_output_meta *meta)
{
    if (!path || !meta) return SPHGPU_E_ARG;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return SPHGPU_E_ARG;
    std::string s;
    s += "# Energy log for SPH simulation\n";
    s += std::string("# Units: ") + meta->unit_type + "\n";
    s += std::string("# Time unit: ") + meta->time_unit + "\n";
    s += std::string("# Energy unit: ") + meta->energy_unit + "\n";
    s += "# Columns: time  kinetic  thermal  potential  total\n";
    s += "#\n";
    bool ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : SPHGPU_E_ARG;
}

int sphgpu_energy_file_write(const char *path, double time, double kinetic, double thermal, double potential,
                             int32_t precision)
{
    if (!path || precision < 0 || precision > 40) return SPHGPU_E_ARG;
    const double total = kinetic + thermal + potential;
    std::string s;
    const double v[5] = {time, kinetic, thermal, potential, total};
    for (int k = 0; k < 5; ++k) {
        put_double(s, v[k], precision);
        s += (k == 4) ? "\n" : "  ";
    }
    FILE *fp = std::fopen(path, "ab");
    if (!fp) return SPHGPU_E_ARG;
    bool ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : SPHGPU_E_ARG;
}

int sphgpu_energy_file_append(sphgpu_ctx *ctx, const char *path, int32_t precision)
{
    if (!ctx || !path) return SPHGPU_E_ARG;
    double t = 0.0, e[7];
    int rc;
    if ((rc = sphgpu_get_time(ctx, &t))) return rc;
    if ((rc = sphgpu_energy_momentum(ctx, e))) return rc;
    return sphgpu_energy_file_write(path, t, e[0], e[1], e[2], precision);
}

} // extern "C"

// ---- legacy VTK --------------------------------------------------------------
// VTKWriter (src/writers/vtk_writer.cpp:17-331): an UNSTRUCTURED_GRID of one
// VTK_VERTEX cell per particle, positions and fields as float.  BINARY writes
// big-endian 4-byte values with no line break after a block, so the next
// keyword line follows the last value directly; ASCII prints each double with
// the stream's default format ("%g", 6 significant digits), absent axes as
// the literal "0.0", and a blank line after every block but CELLS and
// CELL_TYPES.
namespace {

void put_be32(std::string &s, uint32_t v)
{
    const char b[4] = {(char)(v >> 24), (char)(v >> 16), (char)(v >> 8), (char)v};
    s.append(b, 4);
}

void put_be_float(std::string &s, double v)
{
    const float x = (float)v;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    put_be32(s, u);
}

void put_g(std::string &s, double v)
{
    char buf[40];
    const int k = std::snprintf(buf, sizeof(buf), "%g", v);
    s.append(buf, (size_t)k);
}

// one block of per-particle values, formatted by host threads
template <typename F> void vtk_block(std::string &out, long long n, F row)
{
    const int t = worker_count(n);
    std::vector<std::string> part((size_t)std::max(t, 1));
    parallel_ranges(n, [&](int k, long long a, long long b) {
        std::string &s = part[(size_t)k];
        for (long long i = a; i < b; ++i) row(s, i);
    });
    for (const auto &x : part) out += x;
}

} // namespace

extern "C" {

int sphgpu_write_vtk_fields(const char *path, const sphgpu_output_meta *meta, int32_t dim, int32_t n,
                            const double *f, int32_t binary)
{
    if (!path || !meta || n < 0 || (n > 0 && !f) || dim < 1 || dim > 3) return SPHGPU_E_ARG;
    const bool bin = binary != 0;
    std::string s;
    // VTKWriter::write_header (vtk_writer.cpp:40-52)
    s += "# vtk DataFile Version 3.0\n";
    s += std::string("SPH Simulation - ") + meta->simulation_name + "\n";
    s += bin ? "BINARY\n" : "ASCII\n";
    s += "DATASET UNSTRUCTURED_GRID\n";
    const size_t N = (size_t)n;
    auto fld = [&](int row, long long i) { return f[(size_t)row * N + (size_t)i]; };
    auto vec3 = [&](int base) {
        return [&, base](std::string &o, long long i) {
            for (int a = 0; a < 3; ++a) {
                if (bin) {
                    put_be_float(o, a < dim ? fld(base + a, i) : 0.0);
                } else if (a < dim) {
                    if (a) o += ' ';
                    put_g(o, fld(base + a, i));
                } else {
                    o += " 0.0";
                }
            }
            if (!bin) o += '\n';
        };
    };
    auto scalar = [&](int row) {
        return [&, row](std::string &o, long long i) {
            if (bin) put_be_float(o, fld(row, i));
            else { put_g(o, fld(row, i)); o += '\n'; }
        };
    };
    bool ok = n > 0;  // VTKWriter::write_particles refuses an empty set
    if (ok) {
        // write_points (vtk_writer.cpp:91-126)
        s += "POINTS " + std::to_string(n) + " float\n";
        vtk_block(s, n, vec3(SPHGPU_F_POS));
        if (!bin) s += '\n';
        // cells and cell types (vtk_writer.cpp:67-85)
        s += "CELLS " + std::to_string(n) + " " + std::to_string(2ll * n) + "\n";
        vtk_block(s, n, [&](std::string &o, long long i) {
            if (bin) { put_be32(o, 1u); put_be32(o, (uint32_t)(int32_t)i); }
            else { o += "1 "; put_int(o, i); o += '\n'; }
        });
        s += "CELL_TYPES " + std::to_string(n) + "\n";
        vtk_block(s, n, [&](std::string &o, long long) {
            if (bin) put_be32(o, 1u);
            else o += "1\n";
        });
        // write_point_data (vtk_writer.cpp:128-223)
        s += "POINT_DATA " + std::to_string(n) + "\n";
        const std::pair<const char *, int> scalars[] = {
            {"density", SPHGPU_F_DENS}, {"pressure", SPHGPU_F_PRES}, {"mass", SPHGPU_F_MASS},
            {"internal_energy", SPHGPU_F_ENE}, {"smoothing_length", SPHGPU_F_SML}, {"sound_speed", SPHGPU_F_SOUND}};
        for (const auto &sc : scalars) {
            s += std::string("SCALARS ") + sc.first + " float 1\n";
            s += "LOOKUP_TABLE default\n";
            vtk_block(s, n, scalar(sc.second));
            if (!bin) s += '\n';
        }
        s += "SCALARS particle_id int 1\n";
        s += "LOOKUP_TABLE default\n";
        vtk_block(s, n, [&](std::string &o, long long i) {
            const int32_t id = (int32_t)fld(SPHGPU_F_ID, i);
            if (bin) put_be32(o, (uint32_t)id);
            else { put_int(o, id); o += '\n'; }
        });
        if (!bin) s += '\n';
        // write_vectors (vtk_writer.cpp:225-293)
        s += "VECTORS velocity float\n";
        vtk_block(s, n, vec3(SPHGPU_F_VEL));
        if (!bin) s += '\n';
        s += "VECTORS acceleration float\n";
        vtk_block(s, n, vec3(SPHGPU_F_ACC));
        if (!bin) s += '\n';
    }
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return SPHGPU_E_ARG;
    bool wr = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    wr = (std::fclose(fp) == 0) && wr;
    return (ok && wr) ? 0 : SPHGPU_E_ARG;
}

int sphgpu_write_snapshot_vtk(sphgpu_ctx *ctx, const char *path, const sphgpu_output_meta *meta, int32_t binary)
{
    if (!ctx || !path || !meta) return SPHGPU_E_ARG;
    int32_t n = 0;
    sphgpu_params p;
    int rc;
    if ((rc = sphgpu_get_params(ctx, &p, &n))) return rc;
    std::vector<double> f((size_t)SPHGPU_NFIELDS * n);
    if ((rc = sphgpu_download(ctx, f.data(), SPHGPU_MASK_ALL))) return rc;
    return sphgpu_write_vtk_fields(path, meta, p.dim, n, f.data(), binary);
}

} // extern "C"
