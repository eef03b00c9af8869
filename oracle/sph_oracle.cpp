/*
 * sph_oracle.cpp -- CPU restatement of the reference SPHCODE per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against; it is loaded only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product path.
 *
 * It restates, in FP64 and with the reference's operation order, the
 * reference modules listed in each function's comment (paths relative to the
 * reference repository root).  Each floating-point expression keeps the
 * association of the reference expression it restates, so that on identical
 * inputs, compiled without FP contraction, the results are bit-identical to
 * the reference's own modules (pinned by tests/test_oracle_vs_reference.py
 * against oracle/_ref and by the committed golden fixtures).
 *
 * Deliberate differences from the reference:
 *  - node pool exhaustion is reported when the 5N pool is full; the reference
 *    checks `remaind < 0` before taking a node, so it can take one node past
 *    the end of its pool (src/bhtree.cpp:222-228);
 *  - energy sums are serial (the reference uses an OpenMP reduction whose
 *    order depends on the thread count, src/solver.cpp:1159);
 *  - total momentum (sum m v) is added; the reference computes none.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "../include/sphgpu.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef double real;

/* include/defines.hpp:28-35 */
static inline real sqr(real x) { return x * x; }
static inline real pow3(real x) { return x * x * x; }
static inline real pow5(real x) { return x * x * x * x * x; }
static inline real pow6(real x) { return x * x * x * x * x * x; }
static const int neighbor_list_size = 1000;

template <int D> struct Vec {
    real a[D];
    Vec() { for (int k = 0; k < D; ++k) a[k] = 0.0; }
    real &operator[](int k) { return a[k]; }
    const real &operator[](int k) const { return a[k]; }
};

/* include/vector_type.hpp:192-220 (inner_product, abs2, std::abs) */
template <int D> static inline real inner(const Vec<D> &x, const Vec<D> &y)
{
    real s = x[0] * y[0];
    for (int k = 1; k < D; ++k) s = s + x[k] * y[k];
    return s;
}
template <int D> static inline real norm(const Vec<D> &x) { return std::sqrt(inner(x, x)); }

/* include/kernel/kernel_function.hpp:8-16 */
template <int D> static inline real powh(real h)
{
    if (D == 1) return h;
    if (D == 2) return h * h;
    return h * h * h;
}

/* include/kernel/cubic_spline.hpp:13-52, wendland_kernel.hpp:15-48,
 * gaussian_kernel.hpp:13-53.  dw returns the scalar c with grad W = r_ij * c. */
template <int D> struct Kernel {
    int type;
    real sigma_cubic() const
    {
        if (D == 1) return 2.0 / 3.0;
        if (D == 2) return 10.0 / (7.0 * M_PI);
        return 1.0 / M_PI;
    }
    real sigma_c4() const
    {
        if (D == 1) return 0.0;
        if (D == 2) return 9.0 / M_PI;
        return 495. / (32 * M_PI);
    }
    real sigma_gauss() const
    {
        if (D == 1) return 0.564189583547756;
        if (D == 2) return 1.0 / M_PI;
        return 0.179587122125167;
    }
    real w(real r, real h) const
    {
        if (type == SPHGPU_CUBIC_SPLINE) {
            const real h_ = h * 0.5;
            const real q = r / h_;
            return sigma_cubic() / powh<D>(h_) *
                   (0.25 * pow3(0.5 * (2.0 - q + std::abs(2.0 - q))) - pow3(0.5 * (1.0 - q + std::abs(1.0 - q))));
        } else if (type == SPHGPU_WENDLAND) {
            const real q = r / h;
            return sigma_c4() / powh<D>(h) * pow6(0.5 * (1.0 - q + std::abs(1.0 - q))) * (1.0 + 6.0 * q + 35.0 / 3.0 * q * q);
        } else {
            const real q = r / h;
            const real q2 = q * q;
            return sigma_gauss() / powh<D>(h) * std::exp(-q2);
        }
    }
    /* grad W = r_ij * dwc(r, h) */
    real dwc(real r, real h) const
    {
        if (type == SPHGPU_CUBIC_SPLINE) {
            if (r == 0.0) return 0.0;
            const real h_ = h * 0.5;
            const real q = r / h_;
            return -sigma_cubic() / (powh<D>(h_) * h_ * r) *
                   (0.75 * sqr(0.5 * (2.0 - q + std::abs(2.0 - q))) - 3.0 * sqr(0.5 * (1.0 - q + std::abs(1.0 - q))));
        } else if (type == SPHGPU_WENDLAND) {
            const real q = r / h;
            return -56.0 / 3.0 * sigma_c4() / (powh<D>(h) * sqr(h)) * pow5(0.5 * (1.0 - q + std::abs(1.0 - q))) * (1.0 + 5.0 * q);
        } else {
            if (r == 0.0) return 0.0;
            const real q = r / h;
            const real q2 = q * q;
            const real w_val = sigma_gauss() / powh<D>(h) * std::exp(-q2);
            return -2.0 * w_val * q / (h * r);
        }
    }
    real dhw(real r, real h) const
    {
        if (type == SPHGPU_CUBIC_SPLINE) {
            const real h_ = h * 0.5;
            const real q = r / h_;
            return 0.5 * sigma_cubic() / (powh<D>(h_) * h_) *
                   (sqr((std::abs(2.0 - q) + 2.0 - q) * 0.5) * ((3. + D) * 0.25 * q - 0.5 * D) +
                    sqr((std::abs(1.0 - q) + 1.0 - q) * 0.5) * ((-3.0 - D) * q + D));
        } else if (type == SPHGPU_WENDLAND) {
            const double q = r / h;
            return -sigma_c4() / (powh<D>(h) * h * 3.0) * pow5(0.5 * (1.0 - q + std::abs(1.0 - q))) *
                   (3.0 * D + 15.0 * D * q + (-56.0 + 17.0 * D) * q * q - 35.0 * (8.0 + D) * pow3(q));
        } else {
            const real q = r / h;
            const real q2 = q * q;
            const real w_val = sigma_gauss() / powh<D>(h) * std::exp(-q2);
            return w_val * (2.0 * q2 - static_cast<real>(D)) / h;
        }
    }
};

/* include/periodic.hpp:12-73 */
template <int D> struct Periodic {
    bool valid = false;
    Vec<D> mx, mn, range;
    Vec<D> calc_r_ij(const Vec<D> &ri, const Vec<D> &rj) const
    {
        Vec<D> r;
        if (valid) {
            Vec<D> r1, r2, r3;
            for (int k = 0; k < D; ++k) r1[k] = ri[k] - rj[k];
            for (int k = 0; k < D; ++k) r2[k] = r1[k] + range[k];
            for (int k = 0; k < D; ++k) r3[k] = r1[k] - range[k];
            for (int i = 0; i < D; ++i) {
                /* tie order (1,2,3), (2,3,1), (3,1,2): periodic.hpp:42-53 */
                if (std::abs(r1[i]) <= std::abs(r2[i]) && std::abs(r1[i]) <= std::abs(r3[i])) { r[i] = r1[i]; continue; }
                if (std::abs(r2[i]) <= std::abs(r3[i]) && std::abs(r2[i]) <= std::abs(r1[i])) { r[i] = r2[i]; continue; }
                if (std::abs(r3[i]) <= std::abs(r1[i]) && std::abs(r3[i]) <= std::abs(r2[i])) { r[i] = r3[i]; continue; }
                r[i] = r1[i];
            }
        } else {
            for (int k = 0; k < D; ++k) r[k] = ri[k] - rj[k];
        }
        return r;
    }
    void apply(Vec<D> &r) const
    {
        if (!valid) return;
        for (int i = 0; i < D; ++i) {
            if (r[i] < mn[i]) r[i] += range[i];
            else if (r[i] > mx[i]) r[i] -= range[i];
        }
    }
};

template <int D> struct Particle {
    Vec<D> pos, vel, vel_p, acc;
    real mass, dens, pres, ene, ene_p, dene, sml, sound, balsara, alpha, gradh, phi;
    int id, neighbor;
    int next;
};

enum { ERR_NONE = 0 };

template <int D> class Sim {
public:
    static const int NCHILD = D == 1 ? 2 : D == 2 ? 4 : 8;

    /* include/bhtree.hpp:19-60 */
    struct Node {
        int first;
        real mass;
        int num;
        int childs[NCHILD];
        Vec<D> center, m_center;
        real edge;
        int level;
        real kernel_size;
        bool is_leaf;
        void clear()
        {
            first = -1; mass = 0.0; num = 0;
            for (int i = 0; i < NCHILD; ++i) childs[i] = -1;
            center = Vec<D>(); m_center = Vec<D>(); edge = 0.0; level = 0; kernel_size = 0.0; is_leaf = false;
        }
        void root_clear()
        {
            first = -1; mass = 0.0; num = 0; m_center = Vec<D>(); kernel_size = 0.0; is_leaf = false;
        }
    };

    sphgpu_params prm;
    Kernel<D> kernel;
    Periodic<D> periodic;
    std::vector<Particle<D>> p;
    std::vector<Vec<D>> grad_d, grad_p, grad_v[D];
    int num = 0;
    real time = 0.0, dt = 0.0, h_per_v_sig = 0.0;
    bool pre_first = true;
    int err = 0;

    /* tree (src/bhtree.cpp) */
    std::vector<Node> nodes; /* index 0 is the root, 1.. the pool */
    int node_pool = 0;
    int nodes_used = 0;
    bool tree_valid = false;

    explicit Sim(const sphgpu_params &prm_, int n) : prm(prm_)
    {
        kernel.type = prm.kernel;
        num = n;
        p.resize(n);
        grad_d.resize(n); grad_p.resize(n);
        for (int k = 0; k < D; ++k) grad_v[k].resize(n);
        periodic.valid = prm.periodic != 0;
        for (int k = 0; k < D; ++k) {
            periodic.mx[k] = prm.range_max[k];
            periodic.mn[k] = prm.range_min[k];
            periodic.range[k] = periodic.mx[k] - periodic.mn[k];
        }
        /* BHTree::initialize (src/bhtree.cpp:13-41) and resize(N, 5) (:43-57) */
        node_pool = n * 5;
        nodes.resize(node_pool + 1);
        nodes[0].clear();
        nodes[0].level = 1;
        if (periodic.valid) {
            for (int k = 0; k < D; ++k) nodes[0].center[k] = (periodic.mx[k] + periodic.mn[k]) * 0.5;
            real l = 0.0;
            for (int k = 0; k < D; ++k) if (l < periodic.range[k]) l = periodic.range[k];
            nodes[0].edge = l;
        }
        time = 0.0;
        dt = 0.0;
    }

    /* ---------------------------------------------------------------- tree */
    /* BHTree::make (src/bhtree.cpp:59-115) */
    int make_tree()
    {
        Node &root = nodes[0];
        root.root_clear();
        tree_valid = false;
        if (num == 0) return 0;
        if (!periodic.valid) {
            Vec<D> rmin, rmax;
            for (int k = 0; k < D; ++k) { rmin[k] = std::numeric_limits<real>::max(); rmax[k] = std::numeric_limits<real>::lowest(); }
            for (int i = 0; i < num; ++i)
                for (int k = 0; k < D; ++k) {
                    if (rmin[k] > p[i].pos[k]) rmin[k] = p[i].pos[k];
                    if (rmax[k] < p[i].pos[k]) rmax[k] = p[i].pos[k];
                }
            for (int k = 0; k < D; ++k) root.center[k] = (rmax[k] + rmin[k]) * 0.5;
            real l = 0.0;
            for (int k = 0; k < D; ++k) {
                const real range = rmax[k] - rmin[k];
                if (l < range) l = range;
            }
            root.edge = l;
        }
        for (int i = 0; i < num - 1; ++i) p[i].next = i + 1;
        p[num - 1].next = -1;
        root.first = 0;
        nodes_used = 0;
        int e = create_tree(0);
        if (e) return e;
        tree_valid = true;
        return 0;
    }

    /* BHNode::assign (src/bhtree.cpp:206-247) */
    int assign(int node, int pi)
    {
        const Vec<D> &pos = p[pi].pos;
        int index = 0, mask = 1;
        for (int i = 0; i < D; ++i) {
            if (pos[i] > nodes[node].center[i]) index |= mask;
            mask <<= 1;
        }
        int child = nodes[node].childs[index];
        if (child < 0) {
            if (nodes_used >= node_pool) return SPHGPU_E_NO_FREE_NODE;
            child = 1 + nodes_used++;
            nodes[node].childs[index] = child;
            Node &c = nodes[child];
            c.clear();
            c.level = nodes[node].level + 1;
            c.edge = nodes[node].edge * 0.5;
            int a = 1;
            real b = 2.0;
            for (int i = 0; i < D; ++i) {
                c.center[i] = nodes[node].center[i] + ((index & a) * b - 1.0) * nodes[node].edge * 0.25;
                a <<= 1;
                b *= 0.5;
            }
        }
        Node &c = nodes[child];
        c.num++;
        c.mass += p[pi].mass;
        for (int k = 0; k < D; ++k) c.m_center[k] += pos[k] * p[pi].mass;
        p[pi].next = c.first;
        c.first = pi;
        return 0;
    }

    /* BHNode::create_tree (src/bhtree.cpp:154-204) */
    int create_tree(int node)
    {
        for (int i = 0; i < NCHILD; ++i) nodes[node].childs[i] = -1;
        int pp = nodes[node].first;
        do {
            const int pnext = p[pp].next;
            int e = assign(node, pp);
            if (e) return e;
            pp = pnext;
        } while (pp >= 0);

        for (int i = 0; i < NCHILD; ++i) {
            const int child = nodes[node].childs[i];
            if (child >= 0) {
                if (nodes[child].num > prm.leaf_particle_num && nodes[node].level < prm.max_tree_level) {
                    nodes[child].mass = 0.0;
                    nodes[child].m_center = Vec<D>();
                    int e = create_tree(child);
                    if (e) return e;
                } else {
                    nodes[child].is_leaf = true;
                    if (nodes[child].mass > 0.0)
                        for (int k = 0; k < D; ++k) nodes[child].m_center[k] /= nodes[child].mass;
                }
                nodes[node].mass += nodes[child].mass;
                for (int k = 0; k < D; ++k) nodes[node].m_center[k] += nodes[child].m_center[k] * nodes[child].mass;
            }
        }
        if (nodes[node].mass > 0.0)
            for (int k = 0; k < D; ++k) nodes[node].m_center[k] /= nodes[node].mass;
        return 0;
    }

    /* BHNode::set_kernel (src/bhtree.cpp:249-275) */
    real set_kernel(int node)
    {
        real k = 0.0;
        Node &n = nodes[node];
        if (n.is_leaf) {
            for (int q = n.first; q >= 0; q = p[q].next) {
                const real h = p[q].sml;
                if (h > k) k = h;
            }
        } else {
            for (int i = 0; i < NCHILD; ++i) {
                if (n.childs[i] >= 0) {
                    const real h = set_kernel(n.childs[i]);
                    if (h > k) k = h;
                }
            }
        }
        n.kernel_size = k;
        return k;
    }

    /* BHNode::neighbor_search (src/bhtree.cpp:277-316) */
    int node_search(int node, int pi, std::vector<int> &list, int &n_nb, int max_nb, bool is_ij) const
    {
        const Node &n = nodes[node];
        const Vec<D> &r_i = p[pi].pos;
        const real h = is_ij ? std::max(p[pi].sml, n.kernel_size) : p[pi].sml;
        const real h2 = h * h;
        const real l2 = sqr(n.edge * 0.5 + h);
        const Vec<D> d = periodic.calc_r_ij(r_i, n.center);
        real dx2_max = sqr(d[0]);
        for (int i = 1; i < D; ++i) {
            const real dx2 = sqr(d[i]);
            if (dx2 > dx2_max) dx2_max = dx2;
        }
        if (dx2_max <= l2) {
            if (n.is_leaf) {
                for (int q = n.first; q >= 0; q = p[q].next) {
                    const Vec<D> r_ij = periodic.calc_r_ij(r_i, p[q].pos);
                    const real r2 = inner(r_ij, r_ij);
                    if (r2 < h2) {
                        if (n_nb >= max_nb) return SPHGPU_E_NEIGHBOR_OVERFLOW;
                        list[n_nb] = p[q].id;
                        ++n_nb;
                    }
                }
            } else {
                for (int i = 0; i < NCHILD; ++i)
                    if (n.childs[i] >= 0) {
                        int e = node_search(n.childs[i], pi, list, n_nb, max_nb, is_ij);
                        if (e) return e;
                    }
            }
        }
        return 0;
    }

    /* BHTree::neighbor_search (src/bhtree.cpp:122-144); returns count or <0 */
    int neighbor_search(int pi, std::vector<int> &list, bool is_ij) const
    {
        int n_nb = 0;
        const int max_nb = (int)list.size();
        int e = node_search(0, pi, list, n_nb, max_nb, is_ij);
        if (e) return e;
        const Vec<D> &pos_i = p[pi].pos;
        std::sort(list.begin(), list.begin() + n_nb, [&](const int a, const int b) {
            const Vec<D> r_ia = periodic.calc_r_ij(pos_i, p[a].pos);
            const Vec<D> r_ib = periodic.calc_r_ij(pos_i, p[b].pos);
            return inner(r_ia, r_ia) < inner(r_ib, r_ib);
        });
        return n_nb;
    }

    /* Hernquist & Katz (1989) softening, src/bhtree.cpp:319-345 */
    static real grav_f(const real r, const real h)
    {
        const real e = h * 0.5;
        const real u = r / e;
        if (u < 1.0) return (-0.5 * u * u * (1.0 / 3.0 - 3.0 / 20 * u * u + u * u * u / 20) + 1.4) / e;
        else if (u < 2.0) return -1.0 / (15 * r) + (-u * u * (4.0 / 3.0 - u + 0.3 * u * u - u * u * u / 30) + 1.6) / e;
        else return 1 / r;
    }
    static real grav_g(const real r, const real h)
    {
        const real e = h * 0.5;
        const real u = r / e;
        if (u < 1.0) return (4.0 / 3.0 - 1.2 * u * u + 0.5 * u * u * u) / (e * e * e);
        else if (u < 2.0)
            return (-1.0 / 15 + 8.0 / 3 * u * u * u - 3 * u * u * u * u + 1.2 * u * u * u * u * u - u * u * u * u * u * u / 6.0) / (r * r * r);
        else return 1 / (r * r * r);
    }

    /* BHNode::calc_force (src/bhtree.cpp:347-393) */
    void calc_force(int node, int pi, const real theta2, const real G)
    {
        const Node &n = nodes[node];
        Particle<D> &pi_ = p[pi];
        const Vec<D> &r_i = pi_.pos;
        const real l2 = n.edge * n.edge;
        const Vec<D> d = periodic.calc_r_ij(r_i, n.m_center);
        const real d2 = inner(d, d);
        const bool theta_open = (l2 > theta2 * d2);
        const bool distance_open = (d2 < l2);
        if (theta_open || distance_open) {
            if (n.is_leaf) {
                for (int q = n.first; q >= 0; q = p[q].next) {
                    if (p[q].id != pi_.id) {
                        const Vec<D> r_ij = periodic.calc_r_ij(r_i, p[q].pos);
                        const real r = norm(r_ij);
                        pi_.phi -= G * p[q].mass * (grav_f(r, pi_.sml) + grav_f(r, p[q].sml)) * 0.5;
                        const real s = G * p[q].mass * (grav_g(r, pi_.sml) + grav_g(r, p[q].sml)) * 0.5;
                        for (int k = 0; k < D; ++k) pi_.acc[k] -= r_ij[k] * s;
                    }
                }
            } else {
                for (int i = 0; i < NCHILD; ++i)
                    if (n.childs[i] >= 0) calc_force(n.childs[i], pi, theta2, G);
            }
        } else {
            const real r_inv = 1.0 / std::sqrt(d2);
            pi_.phi -= G * n.mass * r_inv;
            const real s = G * n.mass * pow3(r_inv);
            for (int k = 0; k < D; ++k) pi_.acc[k] -= d[k] * s;
        }
    }

    /* ---------------------------------------------------- PreInteraction */
    static constexpr real A_D() { return D == 1 ? 2.0 : D == 2 ? M_PI : 4.0 * M_PI / 3.0; }
    real kernel_ratio() const { return prm.iterative_sml ? 1.2 : 1.0; }
    static inline real powh_(real h) { return D == 1 ? 1 : D == 2 ? h : h * h; }

    /* PreInteraction::newton_raphson (src/pre_interaction.cpp:227-283) and the
     * DISPH number-density version (src/disph/d_pre_interaction.cpp:174-230) */
    real newton_raphson(int pi, const std::vector<int> &list, int n_nb, bool disph, int *iters)
    {
        const Particle<D> &pi_ = p[pi];
        real h_i = pi_.sml / kernel_ratio();
        const real A = A_D();
        const real b = disph ? prm.neighbor_number / A : pi_.mass * prm.neighbor_number / A;
        const real epsilon = 1e-4;
        const int max_iter = 10;
        const Vec<D> &r_i = pi_.pos;
        for (int it = 0; it < max_iter; ++it) {
            const real h_b = h_i;
            real dens = 0.0, ddens = 0.0;
            for (int n = 0; n < n_nb; ++n) {
                const int j = list[n];
                const Vec<D> r_ij = periodic.calc_r_ij(r_i, p[j].pos);
                const real r = norm(r_ij);
                if (r >= h_i) break;
                if (disph) {
                    dens += kernel.w(r, h_i);
                    ddens += kernel.dhw(r, h_i);
                } else {
                    dens += p[j].mass * kernel.w(r, h_i);
                    ddens += p[j].mass * kernel.dhw(r, h_i);
                }
            }
            const real f = dens * powh<D>(h_i) - b;
            const real df = ddens * powh<D>(h_i) + D * dens * powh_(h_i);
            h_i -= f / df;
            if (std::abs(h_i - h_b) < (h_i + h_b) * epsilon) {
                if (iters) *iters = it + 1;
                return h_i;
            }
        }
        if (iters) *iters = -max_iter;
        return pi_.sml / kernel_ratio();
    }

    /* PreInteraction::initial_smoothing (src/pre_interaction.cpp:171-215) */
    int initial_smoothing()
    {
        int e_all = 0;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < num; ++i) {
            Particle<D> &pi_ = p[i];
            std::vector<int> list(prm.neighbor_number * neighbor_list_size);
            const real A = A_D();
            pi_.sml = std::pow(prm.neighbor_number * pi_.mass / (pi_.dens * A), 1.0 / D);
            const int n_nb = neighbor_search(i, list, false);
            if (n_nb < 0) {
#pragma omp critical
                e_all = n_nb;
                continue;
            }
            real dens_i = 0.0;
            for (int n = 0; n < n_nb; ++n) {
                const int j = list[n];
                const Vec<D> r_ij = periodic.calc_r_ij(pi_.pos, p[j].pos);
                const real r = norm(r_ij);
                if (r >= pi_.sml) break;
                dens_i += p[j].mass * kernel.w(r, pi_.sml);
            }
            pi_.dens = dens_i;
        }
        return e_all;
    }

    /* PreInteraction::calculation for ssph (src/pre_interaction.cpp:39-169),
     * disph (src/disph/d_pre_interaction.cpp:21-162) and gsph
     * (src/gsph/g_pre_interaction.cpp:27-144); set_kernel at the end. */
    int pre_interaction()
    {
        if (!tree_valid) return SPHGPU_E_STATE;
        if (pre_first) {
            int e = initial_smoothing();
            if (e) return e;
            pre_first = false;
        }
        const int type = prm.sph_type;
        const real gamma = prm.gamma;
        const int nthreads =
#ifdef _OPENMP
            omp_get_max_threads();
#else
            1;
#endif
        std::vector<real> hpvs(nthreads, std::numeric_limits<real>::max());
        int e_all = 0;
        newton_iters_total = 0;
        newton_fail = 0;

#pragma omp parallel for schedule(static)
        for (int i = 0; i < num; ++i) {
            const int tid =
#ifdef _OPENMP
                omp_get_thread_num();
#else
                0;
#endif
            Particle<D> &pi_ = p[i];
            std::vector<int> list(prm.neighbor_number * neighbor_list_size);
            const real A = A_D();
            pi_.sml = std::pow(prm.neighbor_number * pi_.mass / (pi_.dens * A), 1.0 / D) * kernel_ratio();
            const int n_tmp = neighbor_search(i, list, false);
            if (n_tmp < 0) {
#pragma omp critical
                e_all = n_tmp;
                continue;
            }
            if (prm.iterative_sml) {
                int it = 0;
                /* GDISPH uses the DISPH Newton (src/gdisph/gd_pre_interaction.cpp:214-270) */
                pi_.sml = newton_raphson(i, list, n_tmp, type == SPHGPU_DISPH || type == SPHGPU_GDISPH, &it);
#pragma omp atomic
                newton_iters_total += (it < 0 ? -it : it);
                if (it < 0) {
#pragma omp atomic
                    newton_fail += 1;
                }
            }
            real dens_i = 0.0, dh_dens_i = 0.0;
            real pres_i = 0.0, dh_pres_i = 0.0, n_i = 0.0, dh_n_i = 0.0;
            real v_sig_max = pi_.sound * 2.0;
            const Vec<D> &pos_i = pi_.pos;
            int n_nb = 0;
            for (int n = 0; n < n_tmp; ++n) {
                const int j = list[n];
                const Particle<D> &pj = p[j];
                const Vec<D> r_ij = periodic.calc_r_ij(pos_i, pj.pos);
                const real r = norm(r_ij);
                if (r >= pi_.sml) break;
                ++n_nb;
                if (type == SPHGPU_DISPH || type == SPHGPU_GDISPH) {
                    const real w_ij = kernel.w(r, pi_.sml);
                    const real dhw_ij = kernel.dhw(r, pi_.sml);
                    dens_i += pj.mass * w_ij;
                    n_i += w_ij;
                    pres_i += pj.mass * pj.ene * w_ij;
                    dh_pres_i += pj.mass * pj.ene * dhw_ij;
                    dh_n_i += dhw_ij;
                } else if (type == SPHGPU_SSPH) {
                    dens_i += pj.mass * kernel.w(r, pi_.sml);
                    dh_dens_i += pj.mass * kernel.dhw(r, pi_.sml);
                } else {
                    dens_i += pj.mass * kernel.w(r, pi_.sml);
                }
                if (i != j) {
                    Vec<D> v_ij;
                    for (int k = 0; k < D; ++k) v_ij[k] = pi_.vel[k] - pj.vel[k];
                    const real v_sig = pi_.sound + pj.sound - 3.0 * inner(r_ij, v_ij) / r;
                    if (v_sig > v_sig_max) v_sig_max = v_sig;
                }
            }
            pi_.dens = dens_i;
            if (type == SPHGPU_DISPH || type == SPHGPU_GDISPH) {
                pi_.pres = (gamma - 1.0) * pres_i;
                pi_.gradh = pi_.sml / (D * n_i) * dh_pres_i / (1.0 + pi_.sml / (D * n_i) * dh_n_i);
            } else if (type == SPHGPU_SSPH) {
                pi_.pres = (gamma - 1.0) * dens_i * pi_.ene;
                pi_.gradh = 1.0 / (1.0 + pi_.sml / (D * dens_i) * dh_dens_i);
            } else {
                pi_.pres = (gamma - 1.0) * dens_i * pi_.ene;
            }
            pi_.neighbor = n_nb;
            const real hv = pi_.sml / v_sig_max;
            if (hpvs[tid] > hv) hpvs[tid] = hv;

            if ((type == SPHGPU_GSPH || type == SPHGPU_GDISPH) && prm.gsph_2nd_order) {
                /* MUSCL gradients, src/gsph/g_pre_interaction.cpp:112-136,
                 * src/gdisph/gd_pre_interaction.cpp:125-148 */
                Vec<D> dd, du;
                Vec<D> dv[D];
                for (int n = 0; n < n_nb; ++n) {
                    const int j = list[n];
                    const Particle<D> &pj = p[j];
                    const Vec<D> r_ij = periodic.calc_r_ij(pos_i, pj.pos);
                    const real r = norm(r_ij);
                    const real c = kernel.dwc(r, pi_.sml);
                    Vec<D> dw_ij;
                    for (int k = 0; k < D; ++k) dw_ij[k] = r_ij[k] * c;
                    for (int k = 0; k < D; ++k) dd[k] += dw_ij[k] * pj.mass;
                    const real mdu = pj.mass * (pj.ene - pi_.ene);
                    for (int k = 0; k < D; ++k) du[k] += dw_ij[k] * mdu;
                    for (int a = 0; a < D; ++a) {
                        const real mdv = pj.mass * (pj.vel[a] - pi_.vel[a]);
                        for (int k = 0; k < D; ++k) dv[a][k] += dw_ij[k] * mdv;
                    }
                }
                for (int k = 0; k < D; ++k) grad_d[i][k] = dd[k];
                for (int k = 0; k < D; ++k) grad_p[i][k] = (dd[k] * pi_.ene + du[k]) * (gamma - 1.0);
                const real rho_inv = 1.0 / pi_.dens;
                for (int a = 0; a < D; ++a)
                    for (int k = 0; k < D; ++k) grad_v[a][i][k] = dv[a][k] * rho_inv;
            }
            if (type == SPHGPU_GSPH) continue;

            /* Balsara switch and time-dependent AV,
             * src/pre_interaction.cpp:115-161, src/disph/d_pre_interaction.cpp:106-154;
             * GDISPH (src/gdisph/gd_pre_interaction.cpp:150-197) weights the
             * Balsara sums by mass and divides by density, and its
             * time-dependent AV without Balsara is the DISPH one */
            const bool disph = type == SPHGPU_DISPH;
            const bool disph_av = type == SPHGPU_DISPH || type == SPHGPU_GDISPH;
            if (prm.use_balsara_switch && D != 1) {
                real div_v = 0.0;
                Vec<D> rot3;      /* D == 3 */
                real rot2 = 0.0;  /* D == 2 */
                for (int n = 0; n < n_nb; ++n) {
                    const int j = list[n];
                    const Particle<D> &pj = p[j];
                    const Vec<D> r_ij = periodic.calc_r_ij(pos_i, pj.pos);
                    const real r = norm(r_ij);
                    const real c = kernel.dwc(r, pi_.sml);
                    Vec<D> dw, v_ij;
                    for (int k = 0; k < D; ++k) dw[k] = r_ij[k] * c;
                    for (int k = 0; k < D; ++k) v_ij[k] = pi_.vel[k] - pj.vel[k];
                    const real wgt = disph ? pj.mass * pj.ene : pj.mass;
                    div_v -= wgt * inner(v_ij, dw);
                    if (D == 3) {
                        const real cx = v_ij[1] * dw[2 % D] - v_ij[2 % D] * dw[1];
                        const real cy = v_ij[2 % D] * dw[0] - v_ij[0] * dw[2 % D];
                        const real cz = v_ij[0] * dw[1] - v_ij[1] * dw[0];
                        rot3[0] += cx * wgt;
                        rot3[1 % D] += cy * wgt;
                        rot3[2 % D] += cz * wgt;
                    } else if (D == 2) {
                        rot2 += (v_ij[0] * dw[1 % D] - v_ij[1 % D] * dw[0]) * wgt;
                    }
                }
                real rot_abs;
                if (disph) {
                    const real p_inv = (gamma - 1.0) / pi_.pres;
                    div_v *= p_inv;
                    if (D == 3) { for (int k = 0; k < D; ++k) rot3[k] *= p_inv; rot_abs = norm(rot3); }
                    else { rot2 *= p_inv; rot_abs = std::abs(rot2); }
                } else {
                    div_v /= pi_.dens;
                    if (D == 3) { for (int k = 0; k < D; ++k) rot3[k] /= pi_.dens; rot_abs = norm(rot3); }
                    else { rot2 /= pi_.dens; rot_abs = std::abs(rot2); }
                }
                pi_.balsara = std::abs(div_v) / (std::abs(div_v) + rot_abs + 1e-4 * pi_.sound / pi_.sml);
                if (prm.use_time_dependent_av) {
                    const real tau_inv = prm.epsilon_av * pi_.sound / pi_.sml;
                    const real dalpha = (-(pi_.alpha - prm.alpha_min) * tau_inv +
                                         std::max(-div_v, (real)0.0) * (prm.alpha_max - pi_.alpha)) * dt;
                    pi_.alpha += dalpha;
                }
            } else if (prm.use_time_dependent_av) {
                real div_v = 0.0;
                for (int n = 0; n < n_nb; ++n) {
                    const int j = list[n];
                    const Particle<D> &pj = p[j];
                    const Vec<D> r_ij = periodic.calc_r_ij(pos_i, pj.pos);
                    const real r = norm(r_ij);
                    const real c = kernel.dwc(r, pi_.sml);
                    Vec<D> dw, v_ij;
                    for (int k = 0; k < D; ++k) dw[k] = r_ij[k] * c;
                    for (int k = 0; k < D; ++k) v_ij[k] = pi_.vel[k] - pj.vel[k];
                    const real wgt = disph_av ? pj.mass * pj.ene : pj.mass;
                    div_v -= wgt * inner(v_ij, dw);
                }
                if (disph_av) div_v *= (gamma - 1.0) / pi_.pres;
                else div_v /= pi_.dens;
                const real tau_inv = prm.epsilon_av * pi_.sound / pi_.sml;
                const real s_i = std::max(-div_v, (real)0.0);
                pi_.alpha = (pi_.alpha + dt * tau_inv * prm.alpha_min + s_i * dt * prm.alpha_max) /
                            (1.0 + dt * tau_inv + s_i * dt);
            }
        }
        if (e_all) return e_all;
        real m = hpvs[0];
        for (int t = 1; t < nthreads; ++t) if (m > hpvs[t]) m = hpvs[t];
        h_per_v_sig = m;
        set_kernel(0);
        return 0;
    }

    /* ------------------------------------------------------- FluidForce */
    /* FluidForce::artificial_viscosity (src/fluid_force.cpp:89-106) */
    real artificial_viscosity(const Particle<D> &pi_, const Particle<D> &pj, const Vec<D> &r_ij) const
    {
        Vec<D> v_ij;
        for (int k = 0; k < D; ++k) v_ij[k] = pi_.vel[k] - pj.vel[k];
        const real vr = inner(v_ij, r_ij);
        if (vr < 0) {
            const real alpha = 0.5 * (pi_.alpha + pj.alpha);
            const real balsara = 0.5 * (pi_.balsara + pj.balsara);
            const real w_ij = vr / norm(r_ij);
            const real v_sig = pi_.sound + pj.sound - 3.0 * w_ij;
            const real rho_ij_inv = 2.0 / (pi_.dens + pj.dens);
            return -0.5 * balsara * alpha * v_sig * w_ij * rho_ij_inv;
        }
        return 0;
    }
    /* FluidForce::artificial_conductivity (src/fluid_force.cpp:108-116) */
    real artificial_conductivity(const Particle<D> &pi_, const Particle<D> &pj, const Vec<D> &r_ij, const Vec<D> &dw_ij) const
    {
        real v_sig;
        if (prm.use_gravity) {
            Vec<D> v_ij;
            for (int k = 0; k < D; ++k) v_ij[k] = pi_.vel[k] - pj.vel[k];
            v_sig = std::abs(inner(v_ij, r_ij) / norm(r_ij));
        } else {
            v_sig = std::sqrt(2.0 * std::abs(pi_.pres - pj.pres) / (pi_.dens + pj.dens));
        }
        return prm.alpha_ac * pj.mass * v_sig * (pi_.ene - pj.ene) * inner(dw_ij, r_ij) / norm(r_ij);
    }

    /* van Leer limiter, src/gsph/g_fluid_force.cpp:33-41 */
    static inline real limiter(const real dq1, const real dq2)
    {
        const real dq1dq2 = dq1 * dq2;
        if (dq1dq2 <= 0) return 0.0;
        return 2.0 * dq1dq2 / (dq1 + dq2);
    }

    int fluid_force()
    {
        if (!tree_valid) return SPHGPU_E_STATE;
        const int type = prm.sph_type;
        int e_all = 0;
        force_pairs = 0;
        riemann_iters = 0;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < num; ++i) {
            Particle<D> &pi_ = p[i];
            std::vector<int> list(prm.neighbor_number * neighbor_list_size);
            const int n_nb = neighbor_search(i, list, true);
            if (n_nb < 0) {
#pragma omp critical
                e_all = n_nb;
                continue;
            }
            int64_t pairs = 0, rit = 0;
            if (type == SPHGPU_GSPH) force_gsph(i, list, n_nb, pairs, rit);
            else if (type == SPHGPU_GDISPH) force_gdisph(i, list, n_nb, pairs);
            else force_ssph_disph(i, list, n_nb, type == SPHGPU_DISPH, pairs);
#pragma omp atomic
            force_pairs += pairs;
#pragma omp atomic
            riemann_iters += rit;
        }
        return e_all;
    }

    /* src/fluid_force.cpp:26-86 (ssph) and src/disph/d_fluid_force.cpp:26-86 */
    void force_ssph_disph(int i, const std::vector<int> &list, int n_nb, bool disph, int64_t &pairs)
    {
        Particle<D> &pi_ = p[i];
        const real gamma = prm.gamma;
        const Vec<D> &r_i = pi_.pos;
        const Vec<D> &v_i = pi_.vel;
        const real p_per_rho2_i = pi_.pres / sqr(pi_.dens);
        const real gamma2_u_i = sqr(gamma - 1.0) * pi_.ene;
        const real gamma2_u_per_pres_i = gamma2_u_i / pi_.pres;
        const real m_u_inv = 1.0 / (pi_.mass * pi_.ene);
        const real h_i = pi_.sml;
        const real gradh_i = pi_.gradh;
        Vec<D> acc;
        real dene = 0.0;
        for (int n = 0; n < n_nb; ++n) {
            const int j = list[n];
            const Particle<D> &pj = p[j];
            const Vec<D> r_ij = periodic.calc_r_ij(r_i, pj.pos);
            const real r = norm(r_ij);
            if (r >= std::max(h_i, pj.sml) || r == 0.0) continue;
            ++pairs;
            const real ci = kernel.dwc(r, h_i);
            const real cj = kernel.dwc(r, pj.sml);
            Vec<D> dw_i, dw_j, dw_ij, v_ij;
            for (int k = 0; k < D; ++k) dw_i[k] = r_ij[k] * ci;
            for (int k = 0; k < D; ++k) dw_j[k] = r_ij[k] * cj;
            for (int k = 0; k < D; ++k) dw_ij[k] = (dw_i[k] + dw_j[k]) * 0.5;
            for (int k = 0; k < D; ++k) v_ij[k] = v_i[k] - pj.vel[k];
            const real pi_ij = artificial_viscosity(pi_, pj, r_ij);
            const real dene_ac = prm.use_ac ? artificial_conductivity(pi_, pj, r_ij, dw_ij) : 0.0;
            if (!disph) {
                const real s1 = pj.mass * (p_per_rho2_i * gradh_i + 0.5 * pi_ij);
                const real s2 = pj.mass * (pj.pres / sqr(pj.dens) * pj.gradh + 0.5 * pi_ij);
                for (int k = 0; k < D; ++k) acc[k] -= dw_i[k] * s1 + dw_j[k] * s2;
                dene += pj.mass * p_per_rho2_i * gradh_i * inner(v_ij, dw_i) + 0.5 * pj.mass * pi_ij * inner(v_ij, dw_ij) + dene_ac;
            } else {
                const real f_ij = 1.0 - gradh_i / (pj.mass * pj.ene);
                const real f_ji = 1.0 - pj.gradh * m_u_inv;
                const real u_per_pres_j = pj.ene / pj.pres;
                const real s1 = pj.mass * (gamma2_u_per_pres_i * pj.ene * f_ij + 0.5 * pi_ij);
                const real s2 = pj.mass * (gamma2_u_i * u_per_pres_j * f_ji + 0.5 * pi_ij);
                for (int k = 0; k < D; ++k) acc[k] -= dw_i[k] * s1 + dw_j[k] * s2;
                dene += pj.mass * gamma2_u_per_pres_i * pj.ene * f_ij * inner(v_ij, dw_i) + 0.5 * pj.mass * pi_ij * inner(v_ij, dw_ij) + dene_ac;
            }
        }
        pi_.acc = acc;
        pi_.dene = dene;
    }

    int64_t force_pairs = 0, riemann_iters = 0, newton_iters_total = 0, newton_fail = 0;

    /* GSPH states of the pair (right = i, left = j): MUSCL-reconstructed
     * (src/gsph/g_fluid_force.cpp:103-140) or the particles' own */
    void riemann_states(int i, int j, const Vec<D> &e_ij, real r, real r_inv, real ve_i, real ve_j, real left[4],
                        real right[4]) const
    {
        const Particle<D> &pi_ = p[i];
        const Particle<D> &pj = p[j];
        const real gamma = prm.gamma;
        if (prm.gsph_2nd_order) {
            const real delta_i = 0.5 * (1.0 - pi_.sound * dt * r_inv);
            const real delta_j = 0.5 * (1.0 - pj.sound * dt * r_inv);
            const real dv_ij = ve_i - ve_j;
            Vec<D> dv_i, dv_j;
            for (int k = 0; k < D; ++k) {
                dv_i[k] = inner(grad_v[k][i], e_ij);
                dv_j[k] = inner(grad_v[k][j], e_ij);
            }
            const real dve_i = inner(dv_i, e_ij) * r;
            const real dve_j = inner(dv_j, e_ij) * r;
            right[0] = ve_i - limiter(dv_ij, dve_i) * delta_i;
            left[0] = ve_j + limiter(dv_ij, dve_j) * delta_j;
            const real dd_ij = pi_.dens - pj.dens;
            const real dd_i = inner(grad_d[i], e_ij) * r;
            const real dd_j = inner(grad_d[j], e_ij) * r;
            right[1] = pi_.dens - limiter(dd_ij, dd_i) * delta_i;
            left[1] = pj.dens + limiter(dd_ij, dd_j) * delta_j;
            const real dp_ij = pi_.pres - pj.pres;
            const real dp_i = inner(grad_p[i], e_ij) * r;
            const real dp_j = inner(grad_p[j], e_ij) * r;
            right[2] = pi_.pres - limiter(dp_ij, dp_i) * delta_i;
            left[2] = pj.pres + limiter(dp_ij, dp_j) * delta_j;
            right[3] = std::sqrt(gamma * right[2] / right[1]);
            left[3] = std::sqrt(gamma * left[2] / left[1]);
        } else {
            right[0] = ve_i; right[1] = pi_.dens; right[2] = pi_.pres; right[3] = pi_.sound;
            left[0] = ve_j; left[1] = pj.dens; left[2] = pj.pres; left[3] = pj.sound;
        }
    }

    /* src/gdisph/gd_fluid_force.cpp:40-169: the HLL p* of the (reconstructed)
     * pair states replaces the pressures of the DISPH force, with AV/AC */
    void force_gdisph(int i, const std::vector<int> &list, int n_nb, int64_t &pairs)
    {
        Particle<D> &pi_ = p[i];
        const real gamma = prm.gamma;
        const Vec<D> &r_i = pi_.pos;
        const Vec<D> &v_i = pi_.vel;
        const real h_i = pi_.sml;
        const real gamma2_u_i = sqr(gamma - 1.0) * pi_.ene;
        const real gamma2_u_per_pres_i = gamma2_u_i / pi_.pres;
        const real m_u_inv = 1.0 / (pi_.mass * pi_.ene);
        Vec<D> acc;
        real dene = 0.0;
        for (int n = 0; n < n_nb; ++n) {
            const int j = list[n];
            const Particle<D> &pj = p[j];
            const Vec<D> r_ij = periodic.calc_r_ij(r_i, pj.pos);
            const real r = norm(r_ij);
            if (r >= std::max(h_i, pj.sml) || r == 0.0) continue;
            ++pairs;
            const real r_inv = 1.0 / r;
            Vec<D> e_ij;
            for (int k = 0; k < D; ++k) e_ij[k] = r_ij[k] * r_inv;
            const real ve_i = inner(v_i, e_ij);
            const real ve_j = inner(pj.vel, e_ij);
            real right[4], left[4], pstar, vstar;
            riemann_states(i, j, e_ij, r, r_inv, ve_i, ve_j, left, right);
            hll_solver(left, right, pstar, vstar);
            const real ci = kernel.dwc(r, h_i);
            const real cj = kernel.dwc(r, pj.sml);
            Vec<D> dw_i, dw_j, dw_ij, v_ij;
            for (int k = 0; k < D; ++k) dw_i[k] = r_ij[k] * ci;
            for (int k = 0; k < D; ++k) dw_j[k] = r_ij[k] * cj;
            for (int k = 0; k < D; ++k) dw_ij[k] = (dw_i[k] + dw_j[k]) * 0.5;
            for (int k = 0; k < D; ++k) v_ij[k] = v_i[k] - pj.vel[k];
            const real f_ij = 1.0 - pi_.gradh / (pj.mass * pj.ene);
            const real f_ji = 1.0 - pj.gradh * m_u_inv;
            const real u_per_pres_j = pj.ene / pj.pres;
            const real pi_ij = artificial_viscosity(pi_, pj, r_ij);
            const real dene_ac = prm.use_ac ? artificial_conductivity(pi_, pj, r_ij, dw_ij) : 0.0;
            const real pstar_term_i = pstar / pi_.pres;
            const real pstar_term_j = pstar / pj.pres;
            const real s1 = pj.mass * (gamma2_u_per_pres_i * pj.ene * f_ij * pstar_term_i + 0.5 * pi_ij);
            const real s2 = pj.mass * (gamma2_u_i * u_per_pres_j * f_ji * pstar_term_j + 0.5 * pi_ij);
            for (int k = 0; k < D; ++k) acc[k] -= dw_i[k] * s1 + dw_j[k] * s2;
            dene += pj.mass * gamma2_u_per_pres_i * pj.ene * f_ij * pstar_term_i * inner(v_ij, dw_i) +
                    0.5 * pj.mass * pi_ij * inner(v_ij, dw_ij) + dene_ac;
        }
        pi_.acc = acc;
        pi_.dene = dene;
    }

    /* src/gsph/g_fluid_force.cpp:44-171 */
    void force_gsph(int i, const std::vector<int> &list, int n_nb, int64_t &pairs, int64_t &rit)
    {
        Particle<D> &pi_ = p[i];
        const real gamma = prm.gamma;
        const Vec<D> &r_i = pi_.pos;
        const Vec<D> &v_i = pi_.vel;
        const real h_i = pi_.sml;
        const real rho2_inv_i = 1.0 / sqr(pi_.dens);
        Vec<D> acc;
        real dene = 0.0;
        for (int n = 0; n < n_nb; ++n) {
            const int j = list[n];
            const Particle<D> &pj = p[j];
            const Vec<D> r_ij = periodic.calc_r_ij(r_i, pj.pos);
            const real r = norm(r_ij);
            if (r >= std::max(h_i, pj.sml) || r == 0.0) continue;
            ++pairs;
            const real r_inv = 1.0 / r;
            Vec<D> e_ij;
            for (int k = 0; k < D; ++k) e_ij[k] = r_ij[k] * r_inv;
            const real ve_i = inner(v_i, e_ij);
            const real ve_j = inner(pj.vel, e_ij);
            real vstar, pstar;
            real right[4], left[4];
            if (prm.gsph_2nd_order) {
                const real delta_i = 0.5 * (1.0 - pi_.sound * dt * r_inv);
                const real delta_j = 0.5 * (1.0 - pj.sound * dt * r_inv);
                const real dv_ij = ve_i - ve_j;
                Vec<D> dv_i, dv_j;
                for (int k = 0; k < D; ++k) {
                    dv_i[k] = inner(grad_v[k][i], e_ij);
                    dv_j[k] = inner(grad_v[k][j], e_ij);
                }
                const real dve_i = inner(dv_i, e_ij) * r;
                const real dve_j = inner(dv_j, e_ij) * r;
                right[0] = ve_i - limiter(dv_ij, dve_i) * delta_i;
                left[0] = ve_j + limiter(dv_ij, dve_j) * delta_j;
                const real dd_ij = pi_.dens - pj.dens;
                const real dd_i = inner(grad_d[i], e_ij) * r;
                const real dd_j = inner(grad_d[j], e_ij) * r;
                right[1] = pi_.dens - limiter(dd_ij, dd_i) * delta_i;
                left[1] = pj.dens + limiter(dd_ij, dd_j) * delta_j;
                const real dp_ij = pi_.pres - pj.pres;
                const real dp_i = inner(grad_p[i], e_ij) * r;
                const real dp_j = inner(grad_p[j], e_ij) * r;
                right[2] = pi_.pres - limiter(dp_ij, dp_i) * delta_i;
                left[2] = pj.pres + limiter(dp_ij, dp_j) * delta_j;
                right[3] = std::sqrt(gamma * right[2] / right[1]);
                left[3] = std::sqrt(gamma * left[2] / left[1]);
            } else {
                right[0] = ve_i; right[1] = pi_.dens; right[2] = pi_.pres; right[3] = pi_.sound;
                left[0] = ve_j; left[1] = pj.dens; left[2] = pj.pres; left[3] = pj.sound;
            }
            int it = 0;
            if (prm.riemann_solver == SPHGPU_RIEMANN_ITERATIVE) iterative_solver(gamma, left, right, pstar, vstar, &it);
            else hll_solver(left, right, pstar, vstar);
            rit += it;
            const real ci = kernel.dwc(r, h_i);
            const real cj = kernel.dwc(r, pj.sml);
            Vec<D> dw_i, dw_j, v_ij, f;
            for (int k = 0; k < D; ++k) dw_i[k] = r_ij[k] * ci;
            for (int k = 0; k < D; ++k) dw_j[k] = r_ij[k] * cj;
            for (int k = 0; k < D; ++k) v_ij[k] = e_ij[k] * vstar;
            const real rho2_inv_j = 1.0 / sqr(pj.dens);
            const real s1 = pj.mass * pstar * rho2_inv_i;
            const real s2 = pj.mass * pstar * rho2_inv_j;
            for (int k = 0; k < D; ++k) f[k] = dw_i[k] * s1 + dw_j[k] * s2;
            for (int k = 0; k < D; ++k) acc[k] -= f[k];
            Vec<D> dv;
            for (int k = 0; k < D; ++k) dv[k] = v_ij[k] - v_i[k];
            dene -= inner(f, dv);
        }
        pi_.acc = acc;
        pi_.dene = dene;
    }

    /* FluidForce::hll_solver (src/gsph/g_fluid_force.cpp:173-204) */
    static void hll_solver(const real left[], const real right[], real &pstar, real &vstar)
    {
        const real u_l = left[0], rho_l = left[1], p_l = left[2], c_l = left[3];
        const real u_r = right[0], rho_r = right[1], p_r = right[2], c_r = right[3];
        const real roe_l = std::sqrt(rho_l);
        const real roe_r = std::sqrt(rho_r);
        const real roe_inv = 1.0 / (roe_l + roe_r);
        const real u_t = (roe_l * u_l + roe_r * u_r) * roe_inv;
        const real c_t = (roe_l * c_l + roe_r * c_r) * roe_inv;
        const real s_l = std::min(u_l - c_l, u_t - c_t);
        const real s_r = std::max(u_r + c_r, u_t + c_t);
        const real c1 = rho_l * (s_l - u_l);
        const real c2 = rho_r * (s_r - u_r);
        const real c3 = 1.0 / (c1 - c2);
        const real c4 = p_l - u_l * c1;
        const real c5 = p_r - u_r * c2;
        vstar = (c5 - c4) * c3;
        pstar = (c1 * c5 - c2 * c4) * c3;
    }

    /* FluidForce::iterative_solver (src/gsph/g_fluid_force.cpp:209-304) */
    static void iterative_solver(const real gamma, const real left[], const real right[], real &pstar, real &vstar, int *iters)
    {
        const real u_l = left[0], rho_l = left[1], p_l = left[2];
        const real u_r = right[0], rho_r = right[1], p_r = right[2];
        if (rho_l < 0.0 || rho_r < 0.0 || p_l < 0.0 || p_r < 0.0) {
            pstar = 0.0;
            vstar = 0.0;
            if (iters) *iters = 0;
            return;
        }
        const real gamma2 = 1.0 + gamma;
        const real gamma1 = 0.5 * gamma2 / gamma;
        const real smallp = 1e-25;
        const real V_l = 1.0 / rho_l;
        const real V_r = 1.0 / rho_r;
        const real cl = std::sqrt(gamma * p_l * rho_l);
        const real cr = std::sqrt(gamma * p_r * rho_r);
        real pstar_guess = p_r - p_l - cr * (u_r - u_l);
        pstar_guess = p_l + pstar_guess * cl / (cl + cr);
        pstar_guess = std::max(pstar_guess, smallp);
        const int max_iter = 20;
        const real tol = 1e-6;
        real pstar_iter = pstar_guess;
        int it = 0;
        for (int iter = 0; iter < max_iter; ++iter) {
            const real pstar_old = pstar_iter;
            real w_l = 1.0 + gamma1 * (pstar_iter - p_l) / p_l;
            w_l = cl * std::sqrt(w_l);
            real w_r = 1.0 + gamma1 * (pstar_iter - p_r) / p_r;
            w_r = cr * std::sqrt(w_r);
            real z_l = 4.0 * V_l * w_l * w_l;
            z_l = -z_l * w_l / (z_l - gamma2 * (pstar_iter - p_l));
            real z_r = 4.0 * V_r * w_r * w_r;
            z_r = z_r * w_r / (z_r - gamma2 * (pstar_iter - p_r));
            const real ustar_l = u_l - (pstar_iter - p_l) / w_l;
            const real ustar_r = u_r + (pstar_iter - p_r) / w_r;
            pstar_iter = pstar_iter + (ustar_r - ustar_l) * (z_l * z_r) / (z_r - z_l);
            pstar_iter = std::max(smallp, pstar_iter);
            it = iter + 1;
            if (std::abs(pstar_iter - pstar_old) / pstar_iter < tol) break;
        }
        real w_l = 1.0 + gamma1 * (pstar_iter - p_l) / p_l;
        w_l = cl * std::sqrt(w_l);
        real w_r = 1.0 + gamma1 * (pstar_iter - p_r) / p_r;
        w_r = cr * std::sqrt(w_r);
        const real ustar_l = u_l - (pstar_iter - p_l) / w_l;
        const real ustar_r = u_r + (pstar_iter - p_r) / w_r;
        const real ustar = 0.5 * (ustar_l + ustar_r);
        pstar = pstar_iter;
        vstar = ustar;
        if (iters) *iters = it;
    }

    /* ----------------------------------------------------- GravityForce */
    /* GravityForce::calculation + BHTree::tree_force (src/gravity_force.cpp:52-89,
     * src/bhtree.cpp:146-150) */
    int gravity()
    {
        if (!prm.use_gravity) return 0;
        if (!tree_valid) return SPHGPU_E_STATE;
        const real theta2 = prm.theta * prm.theta;
        const real G = prm.gravity_constant;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < num; ++i) {
            p[i].phi = 0.0;
            calc_force(0, i, theta2, G);
        }
        return 0;
    }

    /* direct O(N^2) gravity, src/gravity_force.cpp:70-84 (EXHAUSTIVE_SEARCH) */
    void gravity_direct()
    {
        const real G = prm.gravity_constant;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < num; ++i) {
            real phi = 0.0;
            Vec<D> force;
            for (int j = 0; j < num; ++j) {
                const Vec<D> r_ij = periodic.calc_r_ij(p[i].pos, p[j].pos);
                const real r = norm(r_ij);
                phi -= G * p[j].mass * (grav_f(r, p[i].sml) + grav_f(r, p[j].sml)) * 0.5;
                const real s = G * p[j].mass * (grav_g(r, p[i].sml) + grav_g(r, p[j].sml)) * 0.5;
                for (int k = 0; k < D; ++k) force[k] -= r_ij[k] * s;
            }
            for (int k = 0; k < D; ++k) p[i].acc[k] += force[k];
            p[i].phi = phi;
        }
    }

    /* --------------------------------------------------------- TimeStep */
    /* TimeStep::calculation (src/timestep.cpp:18-38) */
    void timestep()
    {
        real dt_min = std::numeric_limits<real>::max();
        for (int i = 0; i < num; ++i) {
            const real acc_abs = norm(p[i].acc);
            if (acc_abs > 0.0) {
                const real dt_force_i = prm.cfl_force * std::sqrt(p[i].sml / acc_abs);
                if (dt_force_i < dt_min) dt_min = dt_force_i;
            }
        }
        const real dt_sound_i = prm.cfl_sound * h_per_v_sig;
        dt = std::min(dt_sound_i, dt_min);
    }

    /* Solver::predict, standard branch (src/solver.cpp:1047-1060) */
    void predict()
    {
        const real c_sound = prm.gamma * (prm.gamma - 1.0);
        for (int i = 0; i < num; ++i) {
            Particle<D> &q = p[i];
            for (int k = 0; k < D; ++k) q.vel_p[k] = q.vel[k] + q.acc[k] * (0.5 * dt);
            q.ene_p = q.ene + q.dene * (0.5 * dt);
            for (int k = 0; k < D; ++k) q.pos[k] += q.vel_p[k] * dt;
            for (int k = 0; k < D; ++k) q.vel[k] += q.acc[k] * dt;
            q.ene += q.dene * dt;
            q.sound = std::sqrt(c_sound * q.ene);
            periodic.apply(q.pos);
        }
    }

    /* Solver::correct, standard branch (src/solver.cpp:1106-1111) */
    void correct()
    {
        const real c_sound = prm.gamma * (prm.gamma - 1.0);
        for (int i = 0; i < num; ++i) {
            Particle<D> &q = p[i];
            for (int k = 0; k < D; ++k) q.vel[k] = q.vel_p[k] + q.acc[k] * (0.5 * dt);
            q.ene = q.ene_p + q.dene * (0.5 * dt);
            q.sound = std::sqrt(c_sound * q.ene);
        }
    }

    /* Solver::initialize tail (src/solver.cpp:941-963) */
    int initialize()
    {
        const real c_sound = prm.gamma * (prm.gamma - 1.0);
        for (int i = 0; i < num; ++i) {
            p[i].alpha = prm.av_alpha;
            p[i].balsara = 1.0;
            p[i].sound = std::sqrt(c_sound * p[i].ene);
        }
        int e = make_tree();
        if (e) return e;
        if ((e = pre_interaction())) return e;
        if ((e = fluid_force())) return e;
        return gravity();
    }

    /* Solver::integrate (src/solver.cpp:966-978) + Simulation::update_time */
    int integrate()
    {
        timestep();
        predict();
        int e = make_tree();
        if (e) return e;
        if ((e = pre_interaction())) return e;
        if ((e = fluid_force())) return e;
        if ((e = gravity())) return e;
        correct();
        time += dt;
        return 0;
    }

    /* Solver::compute_total_energies (src/solver.cpp:1149-1175), serial, plus
     * total momentum */
    void energy(double out[7]) const
    {
        real K = 0.0, U = 0.0, W = 0.0;
        real P[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < num; ++i) {
            const real vsq = inner(p[i].vel, p[i].vel);
            K += 0.5 * p[i].mass * vsq;
            U += p[i].mass * p[i].ene;
            if (prm.use_gravity) W += 0.5 * p[i].mass * p[i].phi;
            for (int k = 0; k < D; ++k) P[k] += p[i].mass * p[i].vel[k];
        }
        out[0] = K; out[1] = U; out[2] = W; out[3] = K + U + W;
        out[4] = P[0]; out[5] = P[1]; out[6] = P[2];
    }

    /* ------------------------------------------------ field exchange */
    void set_fields(const double *f, uint64_t mask)
    {
        const int64_t n = num;
#define ROW(fi) (f + (int64_t)(fi) * n)
        for (int i = 0; i < num; ++i) {
            Particle<D> &q = p[i];
            for (int k = 0; k < D; ++k) {
                if (mask & SPHGPU_MASK(SPHGPU_F_POS + k)) q.pos[k] = ROW(SPHGPU_F_POS + k)[i];
                if (mask & SPHGPU_MASK(SPHGPU_F_VEL + k)) q.vel[k] = ROW(SPHGPU_F_VEL + k)[i];
                if (mask & SPHGPU_MASK(SPHGPU_F_VEL_P + k)) q.vel_p[k] = ROW(SPHGPU_F_VEL_P + k)[i];
                if (mask & SPHGPU_MASK(SPHGPU_F_ACC + k)) q.acc[k] = ROW(SPHGPU_F_ACC + k)[i];
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_D + k)) grad_d[i][k] = ROW(SPHGPU_F_GRAD_D + k)[i];
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_P + k)) grad_p[i][k] = ROW(SPHGPU_F_GRAD_P + k)[i];
                for (int a = 0; a < D; ++a)
                    if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_V + 3 * a + k)) grad_v[a][i][k] = ROW(SPHGPU_F_GRAD_V + 3 * a + k)[i];
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
    void get_fields(double *f, uint64_t mask) const
    {
        const int64_t n = num;
        for (int i = 0; i < num; ++i) {
            const Particle<D> &q = p[i];
            for (int k = 0; k < 3; ++k) {
                const bool in = k < D;
                if (mask & SPHGPU_MASK(SPHGPU_F_POS + k)) ROW(SPHGPU_F_POS + k)[i] = in ? q.pos[k % D] : 0.0;
                if (mask & SPHGPU_MASK(SPHGPU_F_VEL + k)) ROW(SPHGPU_F_VEL + k)[i] = in ? q.vel[k % D] : 0.0;
                if (mask & SPHGPU_MASK(SPHGPU_F_VEL_P + k)) ROW(SPHGPU_F_VEL_P + k)[i] = in ? q.vel_p[k % D] : 0.0;
                if (mask & SPHGPU_MASK(SPHGPU_F_ACC + k)) ROW(SPHGPU_F_ACC + k)[i] = in ? q.acc[k % D] : 0.0;
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_D + k)) ROW(SPHGPU_F_GRAD_D + k)[i] = in ? grad_d[i][k % D] : 0.0;
                if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_P + k)) ROW(SPHGPU_F_GRAD_P + k)[i] = in ? grad_p[i][k % D] : 0.0;
                for (int a = 0; a < 3; ++a)
                    if (mask & SPHGPU_MASK(SPHGPU_F_GRAD_V + 3 * a + k))
                        ROW(SPHGPU_F_GRAD_V + 3 * a + k)[i] = (in && a < D) ? grad_v[a % D][i][k % D] : 0.0;
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

    /* pre-order dump of the tree; members in linked-list order */
    void dump_node(int node, std::vector<double> &rec, std::vector<int> &members) const
    {
        const Node &n = nodes[node];
        rec.push_back(n.level);
        rec.push_back(n.num);
        rec.push_back(n.is_leaf ? 1 : 0);
        for (int k = 0; k < 3; ++k) rec.push_back(k < D ? n.center[k % D] : 0.0);
        rec.push_back(n.edge);
        rec.push_back(n.mass);
        for (int k = 0; k < 3; ++k) rec.push_back(k < D ? n.m_center[k % D] : 0.0);
        rec.push_back(n.kernel_size);
        rec.push_back((double)members.size());
        if (n.is_leaf) {
            for (int q = n.first; q >= 0; q = p[q].next) members.push_back(p[q].id);
        } else {
            for (int i = 0; i < NCHILD; ++i)
                if (n.childs[i] >= 0) dump_node(n.childs[i], rec, members);
        }
    }
};

/* ------------------------------------------------------------ C-ABI */
struct Handle {
    int dim;
    void *sim;
};

template <int D> static Sim<D> *S(Handle *h) { return static_cast<Sim<D> *>(h->sim); }

#define DISPATCH(h, expr)                                   \
    do {                                                    \
        if ((h)->dim == 1) { auto *s = S<1>(h); expr; }     \
        else if ((h)->dim == 2) { auto *s = S<2>(h); expr; }\
        else { auto *s = S<3>(h); expr; }                   \
    } while (0)

} // namespace orc

using namespace orc;

extern "C" {

void *orc_create(const sphgpu_params *prm, int n)
{
    if (!prm || n <= 0 || prm->dim < 1 || prm->dim > 3) return nullptr;
    Handle *h = new Handle;
    h->dim = prm->dim;
    if (h->dim == 1) h->sim = new Sim<1>(*prm, n);
    else if (h->dim == 2) h->sim = new Sim<2>(*prm, n);
    else h->sim = new Sim<3>(*prm, n);
    return h;
}

void orc_destroy(void *hp)
{
    Handle *h = (Handle *)hp;
    if (!h) return;
    DISPATCH(h, delete s);
    delete h;
}

void orc_set_fields(void *hp, const double *f, uint64_t mask) { Handle *h = (Handle *)hp; DISPATCH(h, s->set_fields(f, mask)); }
void orc_get_fields(void *hp, double *f, uint64_t mask) { Handle *h = (Handle *)hp; DISPATCH(h, s->get_fields(f, mask)); }
int orc_make_tree(void *hp) { Handle *h = (Handle *)hp; int r = 0; DISPATCH(h, r = s->make_tree()); return r; }
int orc_set_kernel(void *hp) { Handle *h = (Handle *)hp; DISPATCH(h, s->set_kernel(0)); return 0; }
int orc_pre_interaction(void *hp) { Handle *h = (Handle *)hp; int r = 0; DISPATCH(h, r = s->pre_interaction()); return r; }
int orc_fluid_force(void *hp) { Handle *h = (Handle *)hp; int r = 0; DISPATCH(h, r = s->fluid_force()); return r; }
int orc_gravity(void *hp) { Handle *h = (Handle *)hp; int r = 0; DISPATCH(h, r = s->gravity()); return r; }
int orc_gravity_direct(void *hp) { Handle *h = (Handle *)hp; DISPATCH(h, s->gravity_direct()); return 0; }
int orc_timestep(void *hp) { Handle *h = (Handle *)hp; DISPATCH(h, s->timestep()); return 0; }
int orc_predict(void *hp) { Handle *h = (Handle *)hp; DISPATCH(h, s->predict()); return 0; }
int orc_correct(void *hp) { Handle *h = (Handle *)hp; DISPATCH(h, s->correct()); return 0; }
int orc_initialize(void *hp) { Handle *h = (Handle *)hp; int r = 0; DISPATCH(h, r = s->initialize()); return r; }
int orc_integrate(void *hp, int nsteps)
{
    Handle *h = (Handle *)hp;
    int r = 0;
    for (int k = 0; k < nsteps && r == 0; ++k) DISPATCH(h, r = s->integrate());
    return r;
}
void orc_energy(void *hp, double *out) { Handle *h = (Handle *)hp; DISPATCH(h, s->energy(out)); }
double orc_get_dt(void *hp) { Handle *h = (Handle *)hp; double v = 0; DISPATCH(h, v = s->dt); return v; }
void orc_set_dt(void *hp, double dt) { Handle *h = (Handle *)hp; DISPATCH(h, s->dt = dt); }
double orc_get_time(void *hp) { Handle *h = (Handle *)hp; double v = 0; DISPATCH(h, v = s->time); return v; }
double orc_get_h_per_v_sig(void *hp) { Handle *h = (Handle *)hp; double v = 0; DISPATCH(h, v = s->h_per_v_sig); return v; }
void orc_set_pre_first(void *hp, int first) { Handle *h = (Handle *)hp; DISPATCH(h, s->pre_first = first != 0); }
/* counters: [force_pairs, riemann_iters, newton_iters, newton_fail] */
void orc_counters(void *hp, int64_t *out)
{
    Handle *h = (Handle *)hp;
    DISPATCH(h, { out[0] = s->force_pairs; out[1] = s->riemann_iters; out[2] = s->newton_iters_total; out[3] = s->newton_fail; });
}

/* Neighbour list of particle i as BHTree::neighbor_search returns it
 * (sorted by distance); the tree must be built and set_kernel done. */
int orc_neighbor_search(void *hp, int i, int is_ij, int32_t *out, int cap)
{
    Handle *h = (Handle *)hp;
    int r = 0;
    DISPATCH(h, {
        std::vector<int> list(s->prm.neighbor_number * neighbor_list_size);
        r = s->neighbor_search(i, list, is_ij != 0);
        for (int k = 0; k < r && k < cap; ++k) out[k] = list[k];
    });
    return r;
}

/* Tree dump: rec rows of 13 doubles per node (level, num, is_leaf, centre[3],
 * edge, mass, com[3], kernel_size, first-member offset); returns node count.
 * Call with NULL buffers to get the sizes. */
int64_t orc_tree_dump(void *hp, double *rec, int32_t *members, int64_t *n_members)
{
    Handle *h = (Handle *)hp;
    std::vector<double> r;
    std::vector<int> m;
    DISPATCH(h, s->dump_node(0, r, m));
    if (rec) std::memcpy(rec, r.data(), r.size() * sizeof(double));
    if (members) for (size_t k = 0; k < m.size(); ++k) members[k] = m[k];
    if (n_members) *n_members = (int64_t)m.size();
    return (int64_t)(r.size() / 13);
}

/* One Riemann problem with either reference solver; returns the iteration
 * count for the iterative solver. */
int orc_riemann(int solver, double gamma, const double *left, const double *right, double *pstar, double *vstar)
{
    int it = 0;
    if (solver == SPHGPU_RIEMANN_ITERATIVE) Sim<1>::iterative_solver(gamma, left, right, *pstar, *vstar, &it);
    else Sim<1>::hll_solver(left, right, *pstar, *vstar);
    return it;
}

/* Kernel function values for closed-form checks: out = [w, dwc, dhw]. */
void orc_kernel_eval(int dim, int kernel, double r, double h, double *out)
{
    if (dim == 1) { Kernel<1> k; k.type = kernel; out[0] = k.w(r, h); out[1] = k.dwc(r, h); out[2] = k.dhw(r, h); }
    else if (dim == 2) { Kernel<2> k; k.type = kernel; out[0] = k.w(r, h); out[1] = k.dwc(r, h); out[2] = k.dhw(r, h); }
    else { Kernel<3> k; k.type = kernel; out[0] = k.w(r, h); out[1] = k.dwc(r, h); out[2] = k.dhw(r, h); }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

} /* extern "C" */
