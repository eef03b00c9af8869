// csv_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the CPU checker of the snapshot
// and energy file formats; nothing in the product links it).
//
// A single-threaded restatement, with the reference's own stream machinery,
// of:
//   CSVWriter::open / write_header / write_column_names / write_particles
//       (src/writers/csv_writer.cpp:17-165): std::ofstream with
//       std::scientific and std::setprecision(p), values streamed one by one;
//   CSVWriter::read_metadata / read_particles (csv_writer.cpp:174-371):
//       std::getline over 50 header lines, "# Key: Value" parsing with
//       std::stoi / std::stod, then 51 skipped lines and one std::getline
//       split on ',' per row, 21 columns, std::stoi / std::stod per field;
//   OutputManager::initialize (energy header, src/output_manager.cpp:55-73)
//       and OutputManager::write_energy (:235-252).
// The metadata arrives as the library's sphgpu_output_meta (include/sphgpu.h);
// SPH type and kernel names follow OutputMetadata::get_sph_type_name /
// get_kernel_type_name (src/output_metadata.cpp:47-65).
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../include/sphgpu.h"

namespace {

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

void header(std::ofstream &f, const sphgpu_output_meta &m)
{
    f << "# SPH Simulation Output - CSV Format v" << m.format_version << "\n";
    f << "# Timestamp: " << m.timestamp << "\n";
    f << "# Simulation: " << m.simulation_name << "\n";
    f << "#\n";
    f << "# === Unit System ===\n";
    f << "# Type: " << m.unit_type << "\n";
    f << "# Length: " << m.length_unit << " (" << m.length_to_cgs << " cm)\n";
    f << "# Mass: " << m.mass_unit << " (" << m.mass_to_cgs << " g)\n";
    f << "# Time: " << m.time_unit << " (" << m.time_to_cgs << " s)\n";
    f << "# Velocity: " << m.velocity_unit << " (" << m.velocity_to_cgs << " cm/s)\n";
    f << "#\n";
    f << "# === Simulation State ===\n";
    f << "# Step: " << m.step << "\n";
    f << "# Time (code): " << m.time_code << "\n";
    f << "# Time (physical): " << m.time_physical << " " << m.time_unit << "\n";
    f << "# Particle Count: " << m.particle_count << "\n";
    f << "#\n";
    f << "# === Physics Parameters ===\n";
    f << "# Gamma: " << m.gamma << "\n";
    f << "# G: " << m.gravitational_constant << "\n";
    f << "# Neighbor Number: " << m.neighbor_number << "\n";
    f << "# SPH Type: " << type_name(m.sph_type) << "\n";
    f << "# Kernel: " << kern_name(m.kernel) << "\n";
    f << "# Gravity: " << (m.use_gravity ? "enabled" : "disabled") << "\n";
    f << "# Balsara Switch: " << (m.use_balsara ? "enabled" : "disabled") << "\n";
    f << "# Time-Dependent AV: " << (m.use_time_dependent_av ? "enabled" : "disabled") << "\n";
    f << "#\n";
    f << "# === Energy ===\n";
    f << "# Kinetic: " << m.kinetic_energy << " [code units]\n";
    f << "# Thermal: " << m.thermal_energy << " [code units]\n";
    f << "# Potential: " << m.potential_energy << " [code units]\n";
    f << "# Total: " << m.total_energy << " [code units]\n";
    f << "#\n";
    if (m.is_relaxation) {
        f << "# === Lane-Emden Relaxation ===\n";
        f << "# Relaxation Step: " << m.relaxation_step << " / " << m.relaxation_total_steps << "\n";
        f << "# Accumulated Time: " << m.accumulated_time << "\n";
        f << "# Alpha Scaling: " << m.alpha_scaling << "\n";
        f << "# Central Density: " << m.rho_center << "\n";
        f << "# Polytropic K: " << m.K << "\n";
        f << "# Radius: " << m.R << "\n";
        f << "# Total Mass: " << m.M_total << "\n";
        f << "#\n";
    }
    f << "# === Columns ===\n";
    f << "# id: Particle ID\n";
    f << "# pos_x, pos_y, pos_z: Position [code units]\n";
    f << "# vel_x, vel_y, vel_z: Velocity [code units]\n";
    f << "# acc_x, acc_y, acc_z: Acceleration [code units]\n";
    f << "# mass: Particle mass [code units]\n";
    f << "# dens: Density [code units]\n";
    f << "# pres: Pressure [code units]\n";
    f << "# ene: Specific internal energy [code units]\n";
    f << "# sml: Smoothing length [code units]\n";
    f << "# sound: Sound speed [code units]\n";
    f << "# alpha: Artificial viscosity coefficient\n";
    f << "# balsara: Balsara factor\n";
    f << "# gradh: grad-h term factor\n";
    f << "# phi: Gravitational potential [code units]\n";
    f << "# neighbor: Number of neighbors\n";
    f << "#\n";
    f << "id,";
    f << "pos_x,pos_y,pos_z,";
    f << "vel_x,vel_y,vel_z,";
    f << "acc_x,acc_y,acc_z,";
    f << "mass,dens,pres,ene,sml,sound,";
    f << "alpha,balsara,gradh,phi,neighbor\n";
}

std::string trimmed(std::string s)
{
    s.erase(0, s.find_first_not_of(" \t"));
    s.erase(s.find_last_not_of(" \t") + 1);
    return s;
}

} // namespace

extern "C" {

// snapshot from a [SPHGPU_NFIELDS][n] block (index = id), rows in id order;
// axes beyond dim print as 0
int orc_csv_write(const char *path, const sphgpu_output_meta *m, int dim, int n, const double *fl)
{
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    f << std::scientific << std::setprecision(m->precision);
    header(f, *m);
    auto v = [&](int row, int i) { return fl[(size_t)row * n + i]; };
    auto ax = [&](int base, int a, int i) { return a < dim ? v(base + a, i) : 0.0; };
    for (int i = 0; i < n; ++i) {
        f << (int)v(SPHGPU_F_ID, i) << ",";
        f << ax(SPHGPU_F_POS, 0, i) << "," << ax(SPHGPU_F_POS, 1, i) << "," << ax(SPHGPU_F_POS, 2, i) << ",";
        f << ax(SPHGPU_F_VEL, 0, i) << "," << ax(SPHGPU_F_VEL, 1, i) << "," << ax(SPHGPU_F_VEL, 2, i) << ",";
        f << ax(SPHGPU_F_ACC, 0, i) << "," << ax(SPHGPU_F_ACC, 1, i) << "," << ax(SPHGPU_F_ACC, 2, i) << ",";
        f << v(SPHGPU_F_MASS, i) << ",";
        f << v(SPHGPU_F_DENS, i) << ",";
        f << v(SPHGPU_F_PRES, i) << ",";
        f << v(SPHGPU_F_ENE, i) << ",";
        f << v(SPHGPU_F_SML, i) << ",";
        f << v(SPHGPU_F_SOUND, i) << ",";
        f << v(SPHGPU_F_ALPHA, i) << ",";
        f << v(SPHGPU_F_BALSARA, i) << ",";
        f << v(SPHGPU_F_GRADH, i) << ",";
        f << v(SPHGPU_F_PHI, i) << ",";
        f << (int)v(SPHGPU_F_NEIGHBOR, i) << "\n";
    }
    f.flush();
    return f.good() ? 0 : -1;
}

// header fields as read_metadata parses them into m (fields it does not
// parse are left as given); rows into a [SPHGPU_NFIELDS][n_cap] block at
// column = file order; *rows = the number read.  Returns -1 where the
// reference's read_particles returns false.
int orc_csv_read(const char *path, sphgpu_output_meta *m, int n_cap, double *fl, int *rows)
{
    {
        std::ifstream file(path);
        if (!file.is_open()) return -1;
        std::string line;
        int count = 0;
        while (count < 50 && std::getline(file, line)) {
            ++count;
            if (line.empty() || line[0] != '#') continue;
            const auto colon = line.find(':');
            if (colon == std::string::npos) continue;
            const std::string key = trimmed(line.substr(2, colon - 2));
            const std::string value = trimmed(line.substr(colon + 2));
            auto bracket = [&](double &dst) {
                const auto b = value.find(" [");
                if (b != std::string::npos) dst = std::stod(value.substr(0, b));
            };
            try {
                if (key == "Timestamp") {
                    std::snprintf(m->timestamp, sizeof(m->timestamp), "%s", value.c_str());
                } else if (key == "Simulation") {
                    std::snprintf(m->simulation_name, sizeof(m->simulation_name), "%s", value.c_str());
                } else if (key == "Step") m->step = std::stoi(value);
                else if (key == "Time (code)") m->time_code = std::stod(value);
                else if (key == "Particle Count") m->particle_count = std::stoi(value);
                else if (key == "Gamma") m->gamma = std::stod(value);
                else if (key == "G") m->gravitational_constant = std::stod(value);
                else if (key == "Neighbor Number") m->neighbor_number = std::stoi(value);
                else if (key == "Kinetic") bracket(m->kinetic_energy);
                else if (key == "Thermal") bracket(m->thermal_energy);
                else if (key == "Potential") bracket(m->potential_energy);
                else if (key == "Total") bracket(m->total_energy);
                else if (key == "Relaxation Step") {
                    m->is_relaxation = 1;
                    const auto sl = value.find('/');
                    if (sl != std::string::npos) {
                        m->relaxation_step = std::stoi(value.substr(0, sl));
                        m->relaxation_total_steps = std::stoi(value.substr(sl + 2));
                    }
                } else if (key == "Accumulated Time") m->accumulated_time = std::stod(value);
                else if (key == "Alpha Scaling") m->alpha_scaling = std::stod(value);
                else if (key == "Central Density") m->rho_center = std::stod(value);
                else if (key == "Polytropic K") m->K = std::stod(value);
                else if (key == "Radius") m->R = std::stod(value);
                else if (key == "Total Mass") m->M_total = std::stod(value);
            } catch (const std::exception &) {
                return -1;  // the reference lets these escape read_metadata
            }
        }
    }
    std::ifstream file(path);
    if (!file.is_open()) return -1;
    std::string line;
    for (int i = 0; i < 51; ++i)
        if (!std::getline(file, line)) return -1;
    int r = 0;
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        std::istringstream ss(line);
        std::string field;
        std::vector<std::string> fields;
        while (std::getline(ss, field, ',')) fields.push_back(field);
        if (fields.size() != 21) return -1;
        if (r >= n_cap) return -1;
        try {
            const int rows_of[21] = {SPHGPU_F_ID, SPHGPU_F_POS, SPHGPU_F_POS + 1, SPHGPU_F_POS + 2, SPHGPU_F_VEL,
                                     SPHGPU_F_VEL + 1, SPHGPU_F_VEL + 2, SPHGPU_F_ACC, SPHGPU_F_ACC + 1,
                                     SPHGPU_F_ACC + 2, SPHGPU_F_MASS, SPHGPU_F_DENS, SPHGPU_F_PRES, SPHGPU_F_ENE,
                                     SPHGPU_F_SML, SPHGPU_F_SOUND, SPHGPU_F_ALPHA, SPHGPU_F_BALSARA,
                                     SPHGPU_F_GRADH, SPHGPU_F_PHI, SPHGPU_F_NEIGHBOR};
            for (int c = 0; c < 21; ++c) {
                const double v = (c == 0 || c == 20) ? (double)std::stoi(fields[c]) : std::stod(fields[c]);
                fl[(size_t)rows_of[c] * n_cap + r] = v;
            }
        } catch (const std::exception &) {
            return -1;
        }
        ++r;
    }
    *rows = r;
    return r > 0 ? 0 : -1;
}

int orc_energy_begin(const char *path, const sphgpu_output_meta *m)
{
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    f << std::scientific << std::setprecision(m->precision);
    f << "# Energy log for SPH simulation\n";
    f << "# Units: " << m->unit_type << "\n";
    f << "# Time unit: " << m->time_unit << "\n";
    f << "# Energy unit: " << m->energy_unit << "\n";
    f << "# Columns: time  kinetic  thermal  potential  total\n";
    f << "#\n";
    return f.good() ? 0 : -1;
}

int orc_energy_write(const char *path, double time, double kinetic, double thermal, double potential, int precision)
{
    std::ofstream f(path, std::ios::app);
    if (!f.is_open()) return -1;
    f << std::scientific << std::setprecision(precision);
    const double total = kinetic + thermal + potential;
    f << time << "  " << kinetic << "  " << thermal << "  " << potential << "  " << total << "\n";
    f.flush();
    return f.good() ? 0 : -1;
}

} // extern "C"
