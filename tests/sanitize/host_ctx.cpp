// host_ctx.cpp -- a host-memory context for the sanitizer build.
//
// The snapshot writers and readers (output.cpp, hdf5_io.cpp) reach the
// device state only through six context calls of the C-ABI.  This file
// implements those six over a host field block so that the file-format code,
// context paths included, runs under AddressSanitizer / UBSan on a machine
// without a GPU.  It is test infrastructure: nothing in the library uses it.
#include <cmath>
#include <cstring>
#include <vector>

#include "sphgpu.h"

struct sphgpu_ctx {
    sphgpu_params prm;
    int n;
    double time;
    std::vector<double> f;  // [SPHGPU_NFIELDS][n]
};

extern "C" {

sphgpu_ctx *host_ctx_create(const sphgpu_params *p, int n, const double *fields)
{
    sphgpu_ctx *c = new sphgpu_ctx;
    c->prm = *p;
    c->n = n;
    c->time = 0.0;
    c->f.assign(fields, fields + (size_t)SPHGPU_NFIELDS * n);
    return c;
}

void host_ctx_destroy(sphgpu_ctx *c) { delete c; }

const double *host_ctx_fields(const sphgpu_ctx *c) { return c->f.data(); }

int sphgpu_get_params(const sphgpu_ctx *c, sphgpu_params *p, int32_t *n)
{
    if (!c) return SPHGPU_E_ARG;
    if (p) *p = c->prm;
    if (n) *n = c->n;
    return 0;
}

int sphgpu_get_time(sphgpu_ctx *c, double *t)
{
    if (!c || !t) return SPHGPU_E_ARG;
    *t = c->time;
    return 0;
}

int sphgpu_set_time(sphgpu_ctx *c, double t)
{
    if (!c) return SPHGPU_E_ARG;
    c->time = t;
    return 0;
}

int sphgpu_upload(sphgpu_ctx *c, const double *fields, uint64_t mask)
{
    if (!c || !fields) return SPHGPU_E_ARG;
    for (int r = 0; r < SPHGPU_NFIELDS; ++r)
        if (mask >> r & 1u)
            std::memcpy(&c->f[(size_t)r * c->n], fields + (size_t)r * c->n, sizeof(double) * (size_t)c->n);
    return 0;
}

int sphgpu_download(sphgpu_ctx *c, double *fields, uint64_t mask)
{
    if (!c || !fields) return SPHGPU_E_ARG;
    for (int r = 0; r < SPHGPU_NFIELDS; ++r)
        if (mask >> r & 1u)
            std::memcpy(fields + (size_t)r * c->n, &c->f[(size_t)r * c->n], sizeof(double) * (size_t)c->n);
    return 0;
}

// kinetic, thermal, potential, total, momentum (Solver::compute_total_energies)
int sphgpu_energy_momentum(sphgpu_ctx *c, double out[7])
{
    if (!c || !out) return SPHGPU_E_ARG;
    const int d = c->prm.dim, n = c->n;
    const double *f = c->f.data();
    double k = 0, u = 0, w = 0, p[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const double m = f[(size_t)SPHGPU_F_MASS * n + i];
        double v2 = 0;
        for (int a = 0; a < d; ++a) {
            const double v = f[(size_t)(SPHGPU_F_VEL + a) * n + i];
            v2 += v * v;
            p[a] += m * v;
        }
        k += 0.5 * m * v2;
        u += m * f[(size_t)SPHGPU_F_ENE * n + i];
        w += 0.5 * m * f[(size_t)SPHGPU_F_PHI * n + i];
    }
    out[0] = k;
    out[1] = u;
    out[2] = w;
    out[3] = k + u + w;
    out[4] = p[0];
    out[5] = p[1];
    out[6] = p[2];
    return 0;
}

} // extern "C"
