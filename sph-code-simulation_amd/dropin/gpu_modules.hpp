// gpu_modules.hpp -- drop-in GPU Modules for the reference SPH driver.
//
// These classes subclass the reference's Module types (include/module.hpp:10-14,
// include/pre_interaction.hpp:13-39, include/fluid_force.hpp:11-24,
// include/gravity_force.hpp:10-17, include/timestep.hpp:8-15) and run each
// calculation(sim) through the C-ABI of include/sphgpu.h.  They are built
// together with the reference sources (they include its headers); see the
// integration notes for the two-line change in Solver::initialize.
//
// Each call hands the module's input fields from the Simulation's AoS
// particles to the device, runs the module there and writes the fields it
// produces back, so GPU and CPU modules can be mixed freely.  All GPU modules
// of one Simulation share one device context, which stays resident: a field
// row crosses to the device only when its host values differ from the ones
// the device already holds (the driver's predict/correct and any CPU module
// change rows; a GPU module's own outputs do not need to come back), and the
// device tree is rebuilt only when the positions changed, so FluidForce and
// GravityForce reuse PreInteraction's tree and its derived pair lists.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "fluid_force.hpp"
#include "gravity_force.hpp"
#include "module.hpp"
#include "parameters.hpp"
#include "pre_interaction.hpp"
#include "timestep.hpp"

#include "sphgpu.h"

namespace sph {

class Simulation;

namespace gpu {

// Device state shared by the GPU modules of one Simulation.
class Device {
public:
    static std::shared_ptr<Device> attach(const std::shared_ptr<Simulation> &sim,
                                          const SPHParameters &param);
    ~Device();
    sphgpu_ctx *ctx() const { return m_ctx; }
    // hand the selected fields of the Simulation's particles to the device
    // (only rows whose values changed cross); returns the rows sent
    uint64_t upload(Simulation &sim, uint64_t mask);
    // copy the selected device fields into the Simulation's particles
    void download(Simulation &sim, uint64_t mask);
    // the device tree over the current positions: built unless the last
    // build saw the same positions
    void ensure_tree();
    // the device's kernel_size values over the current smoothing lengths
    void ensure_kernel_size();
    // the kernel sizes were just set on the device (by the pre-interaction)
    void mark_kernel_size() { m_ks = m_tree; }
    void check(int rc, const char *what) const;  // THROW_ERROR on failure
    // rows uploaded / downloaded and tree builds, for measurement
    uint64_t rows_up = 0, rows_down = 0, rows_skipped = 0, tree_builds = 0;

private:
    Device() = default;
    sphgpu_ctx *m_ctx = nullptr;
    int m_num = 0;
    bool m_gsph = false;
    std::vector<double> m_block;  // device values of the synced rows, [SPHGPU_NFIELDS][n]
    std::vector<double> m_stage;  // host values being handed over
    uint64_t m_synced = 0;        // rows of m_block equal to the device's
    bool m_tree = false;          // the device tree matches the device positions
    bool m_ks = false;            // its kernel sizes match the device's h
};

sphgpu_params to_sphgpu_params(const SPHParameters &p);

class PreInteraction : public sph::PreInteraction {
    std::shared_ptr<SPHParameters> m_param;
    std::shared_ptr<Device> m_dev;

public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;
};

class FluidForce : public sph::FluidForce {
    std::shared_ptr<SPHParameters> m_param;
    std::shared_ptr<Device> m_dev;

public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;
};

class GravityForce : public sph::GravityForce {
    std::shared_ptr<SPHParameters> m_param;
    std::shared_ptr<Device> m_dev;

public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;
};

class TimeStep : public sph::TimeStep {
    std::shared_ptr<SPHParameters> m_param;
    std::shared_ptr<Device> m_dev;

public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;
};

} // namespace gpu
} // namespace sph
