// gpu_modules.cpp -- drop-in GPU Modules for the reference SPH driver.
//
// Field sets read and written per module follow the reference modules:
//   PreInteraction  reads pos, vel, mass, dens, ene, sound, alpha, balsara;
//                   writes sml, dens, pres, gradh, neighbor, balsara, alpha,
//                   the GSPH gradient arrays and Simulation::h_per_v_sig
//                   (src/pre_interaction.cpp:39-169, src/gsph/g_pre_interaction.cpp:27-144)
//   FluidForce      reads pos, vel, mass, dens, pres, ene, sml, sound, gradh,
//                   balsara, alpha, gradients, dt; writes acc, dene
//                   (src/fluid_force.cpp:26-86, src/gsph/g_fluid_force.cpp:44-171)
//   GravityForce    reads pos, mass, sml, acc; writes phi, acc (added)
//                   (src/gravity_force.cpp:52-89)
//   TimeStep        reads acc, sml, h_per_v_sig; writes dt (src/timestep.cpp:18-38)
#include "gpu_modules.hpp"

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "bhtree.hpp"
#include "exception.hpp"
#include "particle.hpp"
#include "simulation.hpp"

namespace sph {
namespace gpu {

namespace {

std::mutex g_lock;
std::map<const Simulation *, std::weak_ptr<Device>> g_devices;

constexpr uint64_t M(int f) { return SPHGPU_MASK(f); }
constexpr uint64_t kVec(int f) { return M(f) | M(f + 1) | M(f + 2); }
constexpr uint64_t kGrad = kVec(SPHGPU_F_GRAD_D) | kVec(SPHGPU_F_GRAD_P) | kVec(SPHGPU_F_GRAD_V) |
                           kVec(SPHGPU_F_GRAD_V + 3) | kVec(SPHGPU_F_GRAD_V + 6);

} // namespace

sphgpu_params to_sphgpu_params(const SPHParameters &p)
{
    sphgpu_params q;
    sphgpu_params_default(&q, DIM);
    q.sph_type = p.type == SPHType::SSPH ? SPHGPU_SSPH
               : p.type == SPHType::DISPH ? SPHGPU_DISPH
               : p.type == SPHType::GDISPH ? SPHGPU_GDISPH : SPHGPU_GSPH;
    q.kernel = p.kernel == KernelType::CUBIC_SPLINE ? SPHGPU_CUBIC_SPLINE
             : p.kernel == KernelType::WENDLAND ? SPHGPU_WENDLAND : SPHGPU_GAUSSIAN;
    q.riemann_solver = p.gsph.riemann_solver == RiemannSolverType::ITERATIVE ? SPHGPU_RIEMANN_ITERATIVE
                                                                             : SPHGPU_RIEMANN_HLL;
    q.neighbor_number = p.physics.neighbor_number;
    q.leaf_particle_num = p.tree.leaf_particle_num;
    q.max_tree_level = p.tree.max_level;
    q.iterative_sml = p.iterative_sml;
    q.use_balsara_switch = p.av.use_balsara_switch;
    q.use_time_dependent_av = p.av.use_time_dependent_av;
    q.use_ac = p.ac.is_valid;
    q.periodic = p.periodic.is_valid;
    q.use_gravity = p.gravity.is_valid;
    q.gsph_2nd_order = p.gsph.is_2nd_order;
    q.gamma = p.physics.gamma;
    q.cfl_sound = p.cfl.sound;
    q.cfl_force = p.cfl.force;
    q.av_alpha = p.av.alpha;
    q.alpha_max = p.av.alpha_max;
    q.alpha_min = p.av.alpha_min;
    q.epsilon_av = p.av.epsilon;
    q.alpha_ac = p.ac.alpha;
    q.gravity_constant = p.gravity.constant;
    q.theta = p.gravity.theta;
    for (int k = 0; k < DIM; ++k) {
        q.range_min[k] = p.periodic.range_min[k];
        q.range_max[k] = p.periodic.range_max[k];
    }
    return q;
}

std::shared_ptr<Device> Device::attach(const std::shared_ptr<Simulation> &sim, const SPHParameters &param)
{
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_devices.find(sim.get());
    if (it != g_devices.end()) {
        if (auto d = it->second.lock()) return d;
    }
    std::shared_ptr<Device> d(new Device());
    d->m_num = sim->get_particle_num();
    d->m_gsph = param.type == SPHType::GSPH || param.type == SPHType::GDISPH;
    const sphgpu_params q = to_sphgpu_params(param);
    const int rc = sphgpu_create(&q, d->m_num, &d->m_ctx);
    if (rc != 0) THROW_ERROR("sphgpu_create failed (", rc, ")");
    d->m_block.assign((size_t)SPHGPU_NFIELDS * d->m_num, 0.0);
    d->m_stage.assign((size_t)SPHGPU_NFIELDS * d->m_num, 0.0);
    g_devices[sim.get()] = d;
    return d;
}

Device::~Device()
{
    if (m_ctx) sphgpu_destroy(m_ctx);
}

void Device::check(int rc, const char *what) const
{
    if (rc != 0) THROW_ERROR(what, " failed (", rc, "): ", sphgpu_last_error(m_ctx));
}

uint64_t Device::upload(Simulation &sim, uint64_t mask)
{
    auto &ps = sim.get_particles();
    const long long n = m_num;
    double *b = m_stage.data();
    // only the rows asked for are packed, in one pass over the particles
    const bool vec_pos = (mask & kVec(SPHGPU_F_POS)) != 0, vec_vel = (mask & kVec(SPHGPU_F_VEL)) != 0,
               vec_velp = (mask & kVec(SPHGPU_F_VEL_P)) != 0, vec_acc = (mask & kVec(SPHGPU_F_ACC)) != 0;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        const SPHParticle &p = ps[i];
        for (int k = 0; k < DIM; ++k) {
            if (vec_pos) b[(SPHGPU_F_POS + k) * n + i] = p.pos[k];
            if (vec_vel) b[(SPHGPU_F_VEL + k) * n + i] = p.vel[k];
            if (vec_velp) b[(SPHGPU_F_VEL_P + k) * n + i] = p.vel_p[k];
            if (vec_acc) b[(SPHGPU_F_ACC + k) * n + i] = p.acc[k];
        }
#define PK(F, m) if (mask & M(F)) b[(F) * n + i] = p.m
        PK(SPHGPU_F_MASS, mass); PK(SPHGPU_F_DENS, dens); PK(SPHGPU_F_PRES, pres); PK(SPHGPU_F_ENE, ene);
        PK(SPHGPU_F_ENE_P, ene_p); PK(SPHGPU_F_DENE, dene); PK(SPHGPU_F_SML, sml); PK(SPHGPU_F_SOUND, sound);
        PK(SPHGPU_F_BALSARA, balsara); PK(SPHGPU_F_ALPHA, alpha); PK(SPHGPU_F_GRADH, gradh); PK(SPHGPU_F_PHI, phi);
        PK(SPHGPU_F_NEIGHBOR, neighbor);
#undef PK
    }
    if (m_gsph && (mask & kGrad)) {
        const auto &gd = sim.get_vector_array("grad_density");
        const auto &gp = sim.get_vector_array("grad_pressure");
        for (int a = 0; a < DIM; ++a) {
            const auto &gv = sim.get_vector_array("grad_velocity_" + std::to_string(a));
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < n; ++i)
                for (int k = 0; k < DIM; ++k) b[(SPHGPU_F_GRAD_V + 3 * a + k) * n + i] = gv[i][k];
        }
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < n; ++i)
            for (int k = 0; k < DIM; ++k) {
                b[(SPHGPU_F_GRAD_D + k) * n + i] = gd[i][k];
                b[(SPHGPU_F_GRAD_P + k) * n + i] = gp[i][k];
            }
    } else {
        mask &= ~kGrad;
    }
    // rows whose host values equal the device's stay where they are (axes
    // beyond DIM are never packed and stay as they are)
    for (int k = DIM; k < 3; ++k)
        mask &= ~(M(SPHGPU_F_POS + k) | M(SPHGPU_F_VEL + k) | M(SPHGPU_F_VEL_P + k) | M(SPHGPU_F_ACC + k) |
                  M(SPHGPU_F_GRAD_D + k) | M(SPHGPU_F_GRAD_P + k));
    uint64_t send = 0;
    for (int f = 0; f < SPHGPU_NFIELDS; ++f) {
        if (!(mask & M(f))) continue;
        double *dst = m_block.data() + (size_t)f * n;
        const double *src = b + (size_t)f * n;
        if ((m_synced & M(f)) && std::memcmp(dst, src, sizeof(double) * (size_t)n) == 0) {
            ++rows_skipped;
            continue;
        }
        std::memcpy(dst, src, sizeof(double) * (size_t)n);
        send |= M(f);
    }
    if (send) {
        check(sphgpu_upload(m_ctx, m_block.data(), send), "sphgpu_upload");
        m_synced |= send;
        rows_up += (uint64_t)__builtin_popcountll(send);
        // new positions need a new tree; new smoothing lengths new kernel sizes
        if (send & kVec(SPHGPU_F_POS)) m_tree = false;
        if (send & M(SPHGPU_F_SML)) m_ks = false;
    }
    return send;
}

void Device::ensure_tree()
{
    if (m_tree) return;
    check(sphgpu_build_tree(m_ctx), "sphgpu_build_tree");
    ++tree_builds;
    m_tree = true;
    m_ks = false;
}

void Device::ensure_kernel_size()
{
    ensure_tree();
    if (m_ks) return;
    check(sphgpu_set_kernel(m_ctx), "sphgpu_set_kernel");
    m_ks = true;
}

void Device::download(Simulation &sim, uint64_t mask)
{
    if (!m_gsph) mask &= ~kGrad;
    double *b = m_block.data();
    check(sphgpu_download(m_ctx, b, mask), "sphgpu_download");
    m_synced |= mask;
    rows_down += (uint64_t)__builtin_popcountll(mask);
    auto &ps = sim.get_particles();
    const long long n = m_num;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        SPHParticle &p = ps[i];
        for (int k = 0; k < DIM; ++k) {
            if (mask & M(SPHGPU_F_ACC + k)) p.acc[k] = b[(SPHGPU_F_ACC + k) * n + i];
        }
#define DL(F, m) if (mask & M(F)) p.m = b[F * n + i]
        DL(SPHGPU_F_DENS, dens); DL(SPHGPU_F_PRES, pres); DL(SPHGPU_F_DENE, dene); DL(SPHGPU_F_SML, sml);
        DL(SPHGPU_F_BALSARA, balsara); DL(SPHGPU_F_ALPHA, alpha); DL(SPHGPU_F_GRADH, gradh); DL(SPHGPU_F_PHI, phi);
#undef DL
        if (mask & M(SPHGPU_F_NEIGHBOR)) p.neighbor = (int)b[SPHGPU_F_NEIGHBOR * n + i];
    }
    if (m_gsph && (mask & kGrad)) {
        auto &gd = sim.get_vector_array("grad_density");
        auto &gp = sim.get_vector_array("grad_pressure");
        for (int a = 0; a < DIM; ++a) {
            auto &gv = sim.get_vector_array("grad_velocity_" + std::to_string(a));
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < n; ++i)
                for (int k = 0; k < DIM; ++k) gv[i][k] = b[(SPHGPU_F_GRAD_V + 3 * a + k) * n + i];
        }
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < n; ++i)
            for (int k = 0; k < DIM; ++k) {
                gd[i][k] = b[(SPHGPU_F_GRAD_D + k) * n + i];
                gp[i][k] = b[(SPHGPU_F_GRAD_P + k) * n + i];
            }
    }
}

// ---- PreInteraction ------------------------------------------------------------
void PreInteraction::initialize(std::shared_ptr<SPHParameters> param)
{
    sph::PreInteraction::initialize(param);
    m_param = param;
}

void PreInteraction::calculation(std::shared_ptr<Simulation> sim)
{
    if (!m_dev) m_dev = Device::attach(sim, *m_param);
    const uint64_t in = kVec(SPHGPU_F_POS) | kVec(SPHGPU_F_VEL) | M(SPHGPU_F_MASS) | M(SPHGPU_F_DENS) |
                        M(SPHGPU_F_ENE) | M(SPHGPU_F_SOUND) | M(SPHGPU_F_ALPHA) | M(SPHGPU_F_BALSARA);
    m_dev->upload(*sim, in);
    m_dev->check(sphgpu_set_dt(m_dev->ctx(), sim->get_dt()), "sphgpu_set_dt");
    m_dev->ensure_tree();
    // the pre-interaction ends with set_kernel over its new smoothing lengths
    m_dev->check(sphgpu_pre_interaction(m_dev->ctx()), "sphgpu_pre_interaction");
    const uint64_t out = M(SPHGPU_F_SML) | M(SPHGPU_F_DENS) | M(SPHGPU_F_PRES) | M(SPHGPU_F_GRADH) |
                         M(SPHGPU_F_NEIGHBOR) | M(SPHGPU_F_BALSARA) | M(SPHGPU_F_ALPHA) | kGrad;
    m_dev->download(*sim, out);
    m_dev->mark_kernel_size();
    double hvs = 0.0;
    m_dev->check(sphgpu_get_h_per_v_sig(m_dev->ctx(), &hvs), "sphgpu_get_h_per_v_sig");
    sim->set_h_per_v_sig(hvs);
    // keep the host tree's kernel sizes current for any CPU module that follows
    sim->get_tree()->set_kernel();
    m_first = false;
}

// ---- FluidForce ----------------------------------------------------------------
void FluidForce::initialize(std::shared_ptr<SPHParameters> param)
{
    sph::FluidForce::initialize(param);
    m_param = param;
}

void FluidForce::calculation(std::shared_ptr<Simulation> sim)
{
    if (!m_dev) m_dev = Device::attach(sim, *m_param);
    const uint64_t in = kVec(SPHGPU_F_POS) | kVec(SPHGPU_F_VEL) | M(SPHGPU_F_MASS) | M(SPHGPU_F_DENS) |
                        M(SPHGPU_F_PRES) | M(SPHGPU_F_ENE) | M(SPHGPU_F_SML) | M(SPHGPU_F_SOUND) |
                        M(SPHGPU_F_GRADH) | M(SPHGPU_F_BALSARA) | M(SPHGPU_F_ALPHA) | kGrad;
    m_dev->upload(*sim, in);
    m_dev->check(sphgpu_set_dt(m_dev->ctx(), sim->get_dt()), "sphgpu_set_dt");
    m_dev->ensure_kernel_size();
    m_dev->check(sphgpu_fluid_force(m_dev->ctx()), "sphgpu_fluid_force");
    m_dev->download(*sim, kVec(SPHGPU_F_ACC) | M(SPHGPU_F_DENE));
}

// ---- GravityForce --------------------------------------------------------------
void GravityForce::initialize(std::shared_ptr<SPHParameters> param)
{
    sph::GravityForce::initialize(param);
    m_param = param;
}

void GravityForce::calculation(std::shared_ptr<Simulation> sim)
{
    if (!m_param->gravity.is_valid) return;
    if (!m_dev) m_dev = Device::attach(sim, *m_param);
    const uint64_t in = kVec(SPHGPU_F_POS) | M(SPHGPU_F_MASS) | M(SPHGPU_F_SML) | kVec(SPHGPU_F_ACC);
    m_dev->upload(*sim, in);
    m_dev->ensure_tree();
    m_dev->check(sphgpu_gravity(m_dev->ctx()), "sphgpu_gravity");
    m_dev->download(*sim, kVec(SPHGPU_F_ACC) | M(SPHGPU_F_PHI));
}

// ---- TimeStep ------------------------------------------------------------------
void TimeStep::initialize(std::shared_ptr<SPHParameters> param)
{
    sph::TimeStep::initialize(param);
    m_param = param;
}

void TimeStep::calculation(std::shared_ptr<Simulation> sim)
{
    if (!m_dev) m_dev = Device::attach(sim, *m_param);
    m_dev->upload(*sim, kVec(SPHGPU_F_ACC) | M(SPHGPU_F_SML));
    m_dev->check(sphgpu_set_h_per_v_sig(m_dev->ctx(), sim->get_h_per_v_sig()), "sphgpu_set_h_per_v_sig");
    m_dev->check(sphgpu_timestep(m_dev->ctx()), "sphgpu_timestep");
    double dt = 0.0;
    m_dev->check(sphgpu_get_dt(m_dev->ctx(), &dt), "sphgpu_get_dt");
    sim->set_dt(dt);
}

} // namespace gpu
} // namespace sph
