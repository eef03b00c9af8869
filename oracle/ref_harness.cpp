/*
 * ref_harness.cpp -- C-ABI driver around the reference's own hot-path modules.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
 * the reference sources, unmodified (src/{bhtree,pre_interaction,fluid_force,
 * gravity_force,timestep,simulation}.cpp, src/disph/*.cpp, src/gsph/*.cpp,
 * src/gdisph/*.cpp),
 * once per DIM.  It exposes the same entry points as oracle/sph_oracle.cpp with
 * the prefix ref_ so tests can compare the restatement with the reference.
 *
 * The driver part (initialize tail, predict, correct, energies) cannot be
 * linked from the reference because src/solver.cpp needs Boost, which is not
 * installed; it is restated here from src/solver.cpp:941-1113,1149-1175 in the
 * same way as in the oracle.  The three static data members of sph::Logger
 * (declared at include/logger.hpp:12-14) are defined here because their
 * defining file src/logger.cpp also needs Boost; they only hold the log file
 * stream and take no part in any computation.
 */
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

/* Access to BHTree internals (tree dump) and to the GSPH Riemann solver
 * lambdas, harness-only.  The reference headers use `class` only in class
 * declarations (templates use `typename`), and `struct` vs `class` changes
 * neither layout nor symbol names, so the compiled reference objects link
 * unchanged. */
#define class struct
#define private public
#define protected public
#include "defines.hpp"
#include "parameters.hpp"
#include "particle.hpp"
#include "simulation.hpp"
#include "periodic.hpp"
#include "bhtree.hpp"
#include "logger.hpp"
#include "exception.hpp"
#include "kernel/kernel_function.hpp"
#include "pre_interaction.hpp"
#include "fluid_force.hpp"
#include "gravity_force.hpp"
#include "timestep.hpp"
#include "disph/d_pre_interaction.hpp"
#include "disph/d_fluid_force.hpp"
#include "gsph/g_pre_interaction.hpp"
#include "gsph/g_fluid_force.hpp"
#include "gdisph/gd_pre_interaction.hpp"
#include "gdisph/gd_fluid_force.hpp"
#undef private
#undef protected
#undef class

#include "../include/sphgpu.h"
#ifdef SPHGPU_DROPIN
#include "gpu_modules.hpp"
#endif

namespace sph {
std::string Logger::dir_name;
std::ofstream Logger::log_io;
bool Logger::open_flag = false;
}

namespace {

struct Ref {
    std::shared_ptr<sph::SPHParameters> param;
    std::shared_ptr<sph::Simulation> sim;
    std::shared_ptr<sph::PreInteraction> pre;
    std::shared_ptr<sph::FluidForce> force;
    std::shared_ptr<sph::GravityForce> grav;
    std::shared_ptr<sph::TimeStep> ts;
    sphgpu_params prm;
    int num;
};

int guard(const std::function<void()> &f)
{
    try {
        f();
    } catch (const sph::SPHException &e) {
        std::string m = e.what();
        if (m.find("overflow") != std::string::npos || m.find("exceeds") != std::string::npos)
            return SPHGPU_E_NEIGHBOR_OVERFLOW;
        if (m.find("free node") != std::string::npos) return SPHGPU_E_NO_FREE_NODE;
        return SPHGPU_E_ARG;
    } catch (...) {
        return SPHGPU_E_ARG;
    }
    return 0;
}

} // namespace

extern "C" {

int ref_dim(void) { return DIM; }

void *ref_create(const sphgpu_params *prm, int n)
{
    if (!prm || prm->dim != DIM || n <= 0) return nullptr;
    Ref *r = new Ref;
    r->prm = *prm;
    r->num = n;
    auto p = std::make_shared<sph::SPHParameters>();
    p->time.start = 0.0;
    p->time.end = 1.0;
    p->time.output = 1.0;
    p->time.energy = 1.0;
    p->type = prm->sph_type == SPHGPU_SSPH ? sph::SPHType::SSPH
            : prm->sph_type == SPHGPU_DISPH ? sph::SPHType::DISPH
            : prm->sph_type == SPHGPU_GDISPH ? sph::SPHType::GDISPH : sph::SPHType::GSPH;
    p->cfl.sound = prm->cfl_sound;
    p->cfl.force = prm->cfl_force;
    p->av.alpha = prm->av_alpha;
    p->av.use_balsara_switch = prm->use_balsara_switch != 0;
    p->av.use_time_dependent_av = prm->use_time_dependent_av != 0;
    p->av.alpha_max = prm->alpha_max;
    p->av.alpha_min = prm->alpha_min;
    p->av.epsilon = prm->epsilon_av;
    p->ac.is_valid = prm->use_ac != 0;
    p->ac.alpha = prm->alpha_ac;
    p->tree.max_level = prm->max_tree_level;
    p->tree.leaf_particle_num = prm->leaf_particle_num;
    p->physics.neighbor_number = prm->neighbor_number;
    p->physics.gamma = prm->gamma;
    p->kernel = prm->kernel == SPHGPU_CUBIC_SPLINE ? sph::KernelType::CUBIC_SPLINE
              : prm->kernel == SPHGPU_WENDLAND ? sph::KernelType::WENDLAND : sph::KernelType::GAUSSIAN;
    p->iterative_sml = prm->iterative_sml != 0;
    p->periodic.is_valid = prm->periodic != 0;
    for (int k = 0; k < DIM; ++k) {
        p->periodic.range_max[k] = prm->range_max[k];
        p->periodic.range_min[k] = prm->range_min[k];
    }
    p->gravity.is_valid = prm->use_gravity != 0;
    p->gravity.constant = prm->gravity_constant;
    p->gravity.theta = prm->theta;
    p->gsph.is_2nd_order = prm->gsph_2nd_order != 0;
    p->gsph.riemann_solver = prm->riemann_solver == SPHGPU_RIEMANN_ITERATIVE ? sph::RiemannSolverType::ITERATIVE
                                                                             : sph::RiemannSolverType::HLL;
    r->param = p;
    r->sim = std::make_shared<sph::Simulation>(p);
    r->sim->get_particles().resize(n);
    r->sim->set_particle_num(n);
    for (int i = 0; i < n; ++i) r->sim->get_particles()[i].id = i;
    /* module wiring, src/solver.cpp:686-724 */
#ifdef SPHGPU_DROPIN
    /* the drop-in: the GPU Module adapters in place of the reference modules,
     * the change a maintainer makes to Solver::initialize's module wiring */
    r->ts = std::make_shared<sph::gpu::TimeStep>();
    r->pre = std::make_shared<sph::gpu::PreInteraction>();
    r->force = std::make_shared<sph::gpu::FluidForce>();
    if (p->type == sph::SPHType::GSPH || p->type == sph::SPHType::GDISPH) {
#else
    r->ts = std::make_shared<sph::TimeStep>();
    if (p->type == sph::SPHType::SSPH) {
        r->pre = std::make_shared<sph::PreInteraction>();
        r->force = std::make_shared<sph::FluidForce>();
    } else if (p->type == sph::SPHType::DISPH) {
        r->pre = std::make_shared<sph::disph::PreInteraction>();
        r->force = std::make_shared<sph::disph::FluidForce>();
    } else if (p->type == sph::SPHType::GDISPH) {
        r->pre = std::make_shared<sph::gdisph::PreInteraction>();
        r->force = std::make_shared<sph::gdisph::FluidForce>();
        std::vector<std::string> names = {"grad_density", "grad_pressure", "grad_velocity_0"};
#if DIM >= 2
        names.push_back("grad_velocity_1");
#endif
#if DIM == 3
        names.push_back("grad_velocity_2");
#endif
        r->sim->add_vector_array(names);
    } else {
        r->pre = std::make_shared<sph::gsph::PreInteraction>();
        r->force = std::make_shared<sph::gsph::FluidForce>();
#endif
        std::vector<std::string> names = {"grad_density", "grad_pressure", "grad_velocity_0"};
#if DIM >= 2
        names.push_back("grad_velocity_1");
#endif
#if DIM == 3
        names.push_back("grad_velocity_2");
#endif
        r->sim->add_vector_array(names);
    }
#ifdef SPHGPU_DROPIN
    r->grav = std::make_shared<sph::gpu::GravityForce>();
#else
    r->grav = std::make_shared<sph::GravityForce>();
#endif
    r->ts->initialize(p);
    r->pre->initialize(p);
    r->force->initialize(p);
    r->grav->initialize(p);
    r->sim->get_tree()->resize(n);
    return r;
}

void ref_destroy(void *h) { delete (Ref *)h; }

void ref_set_fields(void *h, const double *f, uint64_t mask)
{
    Ref *r = (Ref *)h;
    const int64_t n = r->num;
    auto &ps = r->sim->get_particles();
#define ROW(fi) (f + (int64_t)(fi) * n)
    const bool gs = r->param->type == sph::SPHType::GSPH || r->param->type == sph::SPHType::GDISPH;
    for (int i = 0; i < r->num; ++i) {
        auto &q = ps[i];
        for (int k = 0; k < DIM; ++k) {
            if (mask & SPHGPU_MASK(SPHGPU_F_POS + k)) q.pos[k] = ROW(SPHGPU_F_POS + k)[i];
            if (mask & SPHGPU_MASK(SPHGPU_F_VEL + k)) q.vel[k] = ROW(SPHGPU_F_VEL + k)[i];
            if (mask & SPHGPU_MASK(SPHGPU_F_VEL_P + k)) q.vel_p[k] = ROW(SPHGPU_F_VEL_P + k)[i];
            if (mask & SPHGPU_MASK(SPHGPU_F_ACC + k)) q.acc[k] = ROW(SPHGPU_F_ACC + k)[i];
            if (gs) {
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_D + k)) r->sim->get_vector_array("grad_density")[i][k] = ROW(SPHGPU_F_GRAD_D + k)[i];
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_P + k)) r->sim->get_vector_array("grad_pressure")[i][k] = ROW(SPHGPU_F_GRAD_P + k)[i];
                for (int a = 0; a < DIM; ++a)
                    if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_V + 3 * a + k))
                        r->sim->get_vector_array("grad_velocity_" + std::to_string(a))[i][k] = ROW(SPHGPU_F_GRAD_V + 3 * a + k)[i];
            }
        }
#define SC(fi, m) if (mask & SPHGPU_MASK(fi)) q.m = ROW(fi)[i]
        SC(SPHGPU_F_MASS, mass); SC(SPHGPU_F_DENS, dens); SC(SPHGPU_F_PRES, pres); SC(SPHGPU_F_ENE, ene);
        SC(SPHGPU_F_ENE_P, ene_p); SC(SPHGPU_F_DENE, dene); SC(SPHGPU_F_SML, sml); SC(SPHGPU_F_SOUND, sound);
        SC(SPHGPU_F_BALSARA, balsara); SC(SPHGPU_F_ALPHA, alpha); SC(SPHGPU_F_GRADH, gradh); SC(SPHGPU_F_PHI, phi);
#undef SC
        if (mask & SPHGPU_MASK(SPHGPU_F_ID)) q.id = (int)ROW(SPHGPU_F_ID)[i];
        if (mask & SPHGPU_MASK(SPHGPU_F_NEIGHBOR)) q.neighbor = (int)ROW(SPHGPU_F_NEIGHBOR)[i];
    }
}

void ref_get_fields(void *h, double *f, uint64_t mask)
{
    Ref *r = (Ref *)h;
    const int64_t n = r->num;
    auto &ps = r->sim->get_particles();
    const bool gs = r->param->type == sph::SPHType::GSPH || r->param->type == sph::SPHType::GDISPH;
    for (int i = 0; i < r->num; ++i) {
        const auto &q = ps[i];
        for (int k = 0; k < 3; ++k) {
            const bool in = k < DIM;
            const int kk = in ? k : 0;
            if (mask & SPHGPU_MASK(SPHGPU_F_POS + k)) ROW(SPHGPU_F_POS + k)[i] = in ? q.pos[kk] : 0.0;
            if (mask & SPHGPU_MASK(SPHGPU_F_VEL + k)) ROW(SPHGPU_F_VEL + k)[i] = in ? q.vel[kk] : 0.0;
            if (mask & SPHGPU_MASK(SPHGPU_F_VEL_P + k)) ROW(SPHGPU_F_VEL_P + k)[i] = in ? q.vel_p[kk] : 0.0;
            if (mask & SPHGPU_MASK(SPHGPU_F_ACC + k)) ROW(SPHGPU_F_ACC + k)[i] = in ? q.acc[kk] : 0.0;
            double gd = 0.0, gp = 0.0;
            if (gs && in) {
                gd = r->sim->get_vector_array("grad_density")[i][kk];
                gp = r->sim->get_vector_array("grad_pressure")[i][kk];
            }
            if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_D + k)) ROW(SPHGPU_F_GRAD_D + k)[i] = gd;
            if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_P + k)) ROW(SPHGPU_F_GRAD_P + k)[i] = gp;
            for (int a = 0; a < 3; ++a) {
                double gv = 0.0;
                if (gs && in && a < DIM) gv = r->sim->get_vector_array("grad_velocity_" + std::to_string(a))[i][kk];
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_V + 3 * a + k)) ROW(SPHGPU_F_GRAD_V + 3 * a + k)[i] = gv;
            }
        }
#define SC(fi, m) if (mask & SPHGPU_MASK(fi)) ROW(fi)[i] = q.m
        SC(SPHGPU_F_MASS, mass); SC(SPHGPU_F_DENS, dens); SC(SPHGPU_F_PRES, pres); SC(SPHGPU_F_ENE, ene);
        SC(SPHGPU_F_ENE_P, ene_p); SC(SPHGPU_F_DENE, dene); SC(SPHGPU_F_SML, sml); SC(SPHGPU_F_SOUND, sound);
        SC(SPHGPU_F_BALSARA, balsara); SC(SPHGPU_F_ALPHA, alpha); SC(SPHGPU_F_GRADH, gradh); SC(SPHGPU_F_PHI, phi);
        SC(SPHGPU_F_ID, id); SC(SPHGPU_F_NEIGHBOR, neighbor);
#undef SC
    }
#undef ROW
}

int ref_make_tree(void *h) { Ref *r = (Ref *)h; return guard([&] { r->sim->make_tree(); }); }
int ref_set_kernel(void *h) { Ref *r = (Ref *)h; return guard([&] { r->sim->get_tree()->set_kernel(); }); }
int ref_pre_interaction(void *h) { Ref *r = (Ref *)h; return guard([&] { r->pre->calculation(r->sim); }); }
int ref_fluid_force(void *h) { Ref *r = (Ref *)h; return guard([&] { r->force->calculation(r->sim); }); }
int ref_gravity(void *h) { Ref *r = (Ref *)h; return guard([&] { r->grav->calculation(r->sim); }); }
int ref_timestep(void *h) { Ref *r = (Ref *)h; return guard([&] { r->ts->calculation(r->sim); }); }

/* Solver::predict / correct, standard branch (src/solver.cpp:1047-1060, 1106-1111) */
int ref_predict(void *h)
{
    Ref *r = (Ref *)h;
    auto &p = r->sim->get_particles();
    auto *periodic = r->sim->get_periodic().get();
    const real dt = r->sim->get_dt();
    const real gamma = r->param->physics.gamma;
    const real c_sound = gamma * (gamma - 1.0);
    for (int i = 0; i < r->num; ++i) {
        p[i].vel_p = p[i].vel + p[i].acc * (0.5 * dt);
        p[i].ene_p = p[i].ene + p[i].dene * (0.5 * dt);
        p[i].pos += p[i].vel_p * dt;
        p[i].vel += p[i].acc * dt;
        p[i].ene += p[i].dene * dt;
        p[i].sound = std::sqrt(c_sound * p[i].ene);
        periodic->apply(p[i].pos);
    }
    return 0;
}

int ref_correct(void *h)
{
    Ref *r = (Ref *)h;
    auto &p = r->sim->get_particles();
    const real dt = r->sim->get_dt();
    const real gamma = r->param->physics.gamma;
    const real c_sound = gamma * (gamma - 1.0);
    for (int i = 0; i < r->num; ++i) {
        p[i].vel = p[i].vel_p + p[i].acc * (0.5 * dt);
        p[i].ene = p[i].ene_p + p[i].dene * (0.5 * dt);
        p[i].sound = std::sqrt(c_sound * p[i].ene);
    }
    return 0;
}

/* Solver::initialize tail (src/solver.cpp:941-963) */
int ref_initialize(void *h)
{
    Ref *r = (Ref *)h;
    auto &p = r->sim->get_particles();
    const real gamma = r->param->physics.gamma;
    const real c_sound = gamma * (gamma - 1.0);
    for (int i = 0; i < r->num; ++i) {
        p[i].alpha = r->param->av.alpha;
        p[i].balsara = 1.0;
        p[i].sound = std::sqrt(c_sound * p[i].ene);
    }
    return guard([&] {
        r->sim->get_tree()->make(p, r->num);
        r->pre->calculation(r->sim);
        r->force->calculation(r->sim);
        r->grav->calculation(r->sim);
    });
}

/* Solver::integrate (src/solver.cpp:966-978) + Simulation::update_time */
int ref_integrate(void *h, int nsteps)
{
    Ref *r = (Ref *)h;
    for (int k = 0; k < nsteps; ++k) {
        int e = guard([&] {
            r->ts->calculation(r->sim);
            ref_predict(h);
            r->sim->make_tree();
            r->pre->calculation(r->sim);
            r->force->calculation(r->sim);
            r->grav->calculation(r->sim);
            ref_correct(h);
            r->sim->update_time();
        });
        if (e) return e;
    }
    return 0;
}

/* Solver::compute_total_energies (src/solver.cpp:1149-1175), serial */
void ref_energy(void *h, double *out)
{
    Ref *r = (Ref *)h;
    const auto &ps = r->sim->get_particles();
    real K = 0.0, U = 0.0, W = 0.0;
    real P[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < r->num; ++i) {
        const auto &p = ps[i];
        const real vsq = inner_product(p.vel, p.vel);
        K += 0.5 * p.mass * vsq;
        U += p.mass * p.ene;
        if (r->param->gravity.is_valid) W += 0.5 * p.mass * p.phi;
        for (int k = 0; k < DIM; ++k) P[k] += p.mass * p.vel[k];
    }
    out[0] = K; out[1] = U; out[2] = W; out[3] = K + U + W;
    out[4] = P[0]; out[5] = P[1]; out[6] = P[2];
}

double ref_get_dt(void *h) { return ((Ref *)h)->sim->get_dt(); }
void ref_set_dt(void *h, double dt) { ((Ref *)h)->sim->set_dt(dt); }
double ref_get_time(void *h) { return ((Ref *)h)->sim->get_time(); }
double ref_get_h_per_v_sig(void *h) { return ((Ref *)h)->sim->get_h_per_v_sig(); }
void ref_set_pre_first(void *h, int first) { ((Ref *)h)->pre->m_first = first != 0; }

int ref_neighbor_search(void *h, int i, int is_ij, int32_t *out, int cap)
{
    Ref *r = (Ref *)h;
    int n = 0;
    int e = guard([&] {
        std::vector<int> list(r->param->physics.neighbor_number * neighbor_list_size);
        n = r->sim->get_tree()->neighbor_search(r->sim->get_particles()[i], list, r->sim->get_particles(), is_ij != 0);
        for (int k = 0; k < n && k < cap; ++k) out[k] = list[k];
    });
    return e ? e : n;
}

static void dump_node(sph::BHTree::BHNode *n, std::vector<double> &rec, std::vector<int> &members)
{
    rec.push_back(n->level);
    rec.push_back(n->num);
    rec.push_back(n->is_leaf ? 1 : 0);
    for (int k = 0; k < 3; ++k) rec.push_back(k < DIM ? n->center[k < DIM ? k : 0] : 0.0);
    rec.push_back(n->edge);
    rec.push_back(n->mass);
    for (int k = 0; k < 3; ++k) rec.push_back(k < DIM ? n->m_center[k < DIM ? k : 0] : 0.0);
    rec.push_back(n->kernel_size);
    rec.push_back((double)members.size());
    if (n->is_leaf) {
        for (auto *p = n->first; p; p = p->next) members.push_back(p->id);
    } else {
        for (int i = 0; i < sph::NCHILD; ++i)
            if (n->childs[i]) dump_node(n->childs[i], rec, members);
    }
}

int64_t ref_tree_dump(void *h, double *rec, int32_t *members, int64_t *n_members)
{
    Ref *r = (Ref *)h;
    std::vector<double> rv;
    std::vector<int> m;
    dump_node(&r->sim->get_tree()->m_root, rv, m);
    if (rec) std::memcpy(rec, rv.data(), rv.size() * sizeof(double));
    if (members) for (size_t k = 0; k < m.size(); ++k) members[k] = m[k];
    if (n_members) *n_members = (int64_t)m.size();
    return (int64_t)(rv.size() / 13);
}

/* Riemann problem through the reference's own solver lambdas
 * (src/gsph/g_fluid_force.cpp:173-304) */
int ref_riemann(int solver, double gamma, const double *left, const double *right, double *pstar, double *vstar)
{
    sph::gsph::FluidForce f;
    f.m_gamma = gamma;
    if (solver == SPHGPU_RIEMANN_ITERATIVE) f.iterative_solver();
    else f.hll_solver();
    real l[4] = {left[0], left[1], left[2], left[3]};
    real rr[4] = {right[0], right[1], right[2], right[3]};
    real ps = 0, vs = 0;
    f.m_solver(l, rr, ps, vs);
    *pstar = ps;
    *vstar = vs;
    return 0;
}

} /* extern "C" */
