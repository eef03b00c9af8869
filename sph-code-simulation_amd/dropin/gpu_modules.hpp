// gpu_modules.hpp -- drop-in GPU Modules for the reference SPH driver.
//
// These classes subclass the reference's Module types (include/module.hpp:10-14,
// include/pre_interaction.hpp:13-39, include/fluid_force.hpp:11-24,
// include/gravity_force.hpp:10-17, include/timestep.hpp:8-15) and run each
// calculation(sim) through the C-ABI of include/sphgpu.h.  They are built
// together with the reference sources (they include its headers); see the
// integration notes for the two-line change in Solver::initialize.
//
// Each call uploads the fields the module reads from the Simulation's AoS
// particles, runs the module on the device and writes the fields it produces
// back, so GPU and CPU modules can be mixed freely.  All GPU modules of one
// Simulation share one device context.
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
    // copy the selected fields between the Simulation's particles and the device
    void upload(Simulation &sim, uint64_t mask);
    void download(Simulation &sim, uint64_t mask);
    void check(int rc, const char *what) const;  // THROW_ERROR on failure

private:
    Device() = default;
    sphgpu_ctx *m_ctx = nullptr;
    int m_num = 0;
    bool m_gsph = false;
    std::vector<double> m_block;  // [SPHGPU_NFIELDS][n]
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
