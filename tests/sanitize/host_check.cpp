// host_check.cpp -- drives the host code under AddressSanitizer and UBSan.
//
//   host_check <dir> <case.bin>...
//
// Each case file holds a parameter block and a field block (written by
// tests/test_host_sanitizers.py from the workload builders).  For every case
// the oracle (oracle/sph_oracle.cpp) initialises and integrates a few steps
// and answers neighbour, tree, Riemann and kernel queries; the resulting
// state then goes through every snapshot format of the library's host code
// (csrc/output.cpp, csrc/meta_json.cpp and, with WITH_HDF5, csrc/hdf5_io.cpp),
// written and read back, including the inputs the readers must reject.  Any
// sanitizer report ends the process with a non-zero status; a failed check
// prints FAIL and exits with status 1.  Test infrastructure only.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "sphgpu.h"
#ifdef WITH_HDF5
#include "sphgpu_hdf5.h"
#endif

extern "C" {
sphgpu_ctx *host_ctx_create(const sphgpu_params *p, int n, const double *fields);
void host_ctx_destroy(sphgpu_ctx *c);
const double *host_ctx_fields(const sphgpu_ctx *c);

void *orc_create(const sphgpu_params *prm, int n);
void orc_destroy(void *h);
void orc_set_fields(void *h, const double *f, uint64_t mask);
void orc_get_fields(void *h, double *f, uint64_t mask);
int orc_initialize(void *h);
int orc_integrate(void *h, int nsteps);
int orc_make_tree(void *h);
int orc_set_kernel(void *h);
void orc_energy(void *h, double *out);
double orc_get_time(void *h);
int orc_neighbor_search(void *h, int i, int is_ij, int32_t *out, int cap);
int64_t orc_tree_dump(void *h, double *rec, int32_t *members, int64_t *n_members);
int orc_riemann(int solver, double gamma, const double *left, const double *right, double *pstar, double *vstar);
void orc_kernel_eval(int dim, int kernel, double r, double h, double *out);
int orc_csv_write(const char *path, const sphgpu_output_meta *m, int dim, int n, const double *fl);
int orc_csv_read(const char *path, sphgpu_output_meta *m, int n_cap, double *fl, int *rows);
}

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);               \
            std::printf(__VA_ARGS__);                                      \
            std::printf("\n");                                             \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

struct Case {
    std::string name;
    sphgpu_params prm;
    int n = 0;
    std::vector<double> f;
};

bool load_case(const std::string &path, Case &c)
{
    std::ifstream in(path, std::ios::binary);
    int32_t magic = 0, psize = 0;
    in.read((char *)&magic, 4);
    in.read((char *)&psize, 4);
    if (!in || magic != 0x53504843 || psize != (int32_t)sizeof(sphgpu_params)) return false;
    in.read((char *)&c.prm, sizeof(sphgpu_params));
    int32_t n = 0;
    in.read((char *)&n, 4);
    if (!in || n <= 0) return false;
    c.n = n;
    c.f.resize((size_t)SPHGPU_NFIELDS * n);
    in.read((char *)c.f.data(), sizeof(double) * c.f.size());
    c.name = path.substr(path.find_last_of('/') + 1);
    return (bool)in;
}

std::string slurp(const std::string &p)
{
    std::ifstream in(p, std::ios::binary);
    return std::string((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

void spit(const std::string &p, const std::string &s)
{
    std::ofstream out(p, std::ios::binary);
    out << s;
}

// the 21 columns the snapshot files carry, for the dimension's axes
std::vector<int> snapshot_rows(int dim)
{
    std::vector<int> r;
    for (int a = 0; a < dim; ++a) r.push_back(SPHGPU_F_POS + a);
    for (int a = 0; a < dim; ++a) r.push_back(SPHGPU_F_VEL + a);
    for (int a = 0; a < dim; ++a) r.push_back(SPHGPU_F_ACC + a);
    for (int k : {SPHGPU_F_MASS, SPHGPU_F_DENS, SPHGPU_F_PRES, SPHGPU_F_ENE, SPHGPU_F_SML, SPHGPU_F_SOUND,
                  SPHGPU_F_ALPHA, SPHGPU_F_BALSARA, SPHGPU_F_GRADH, SPHGPU_F_PHI, SPHGPU_F_ID, SPHGPU_F_NEIGHBOR})
        r.push_back(k);
    return r;
}

bool rows_equal(const std::vector<double> &a, const std::vector<double> &b, int n, const std::vector<int> &rows)
{
    for (int r : rows)
        if (std::memcmp(&a[(size_t)r * n], &b[(size_t)r * n], sizeof(double) * (size_t)n) != 0) return false;
    return true;
}

void run_oracle(Case &c)
{
    void *o = orc_create(&c.prm, c.n);
    CHECK(o, "%s: orc_create", c.name.c_str());
    if (!o) return;
    orc_set_fields(o, c.f.data(), SPHGPU_MASK_ALL);
    CHECK(orc_initialize(o) == 0, "%s: initialize", c.name.c_str());
    double e0[7], e1[7];
    orc_energy(o, e0);
    CHECK(orc_integrate(o, 3) == 0, "%s: integrate", c.name.c_str());
    orc_energy(o, e1);
    CHECK(std::isfinite(e1[3]) && orc_get_time(o) > 0.0, "%s: energy / time", c.name.c_str());
    // neighbour lists of a few particles, both variants, after a fresh tree
    CHECK(orc_make_tree(o) == 0, "%s: make_tree", c.name.c_str());
    orc_set_kernel(o);
    std::vector<int32_t> list(c.n);
    for (int i = 0; i < c.n; i += std::max(1, c.n / 7))
        for (int ij = 0; ij < 2; ++ij) {
            const int k = orc_neighbor_search(o, i, ij, list.data(), c.n);
            CHECK(k >= 0 && k <= c.n, "%s: neighbor_search(%d, %d) = %d", c.name.c_str(), i, ij, k);
        }
    int64_t nm = 0;
    const int64_t nodes = orc_tree_dump(o, nullptr, nullptr, &nm);
    std::vector<double> rec((size_t)nodes * 13);
    std::vector<int32_t> mem((size_t)nm);
    CHECK(orc_tree_dump(o, rec.data(), mem.data(), &nm) == nodes && nm == c.n, "%s: tree_dump", c.name.c_str());
    orc_get_fields(o, c.f.data(), SPHGPU_MASK_ALL);
    orc_destroy(o);
}

void run_solvers()
{
    const double g = 5.0 / 3.0;
    for (int s = 0; s < 2; ++s)
        for (int k = 0; k < 200; ++k) {
            const double pl = std::pow(10.0, -3.0 + 4.0 * ((k * 37) % 101) / 100.0);
            const double pr = std::pow(10.0, -3.0 + 4.0 * ((k * 53) % 97) / 96.0);
            const double l[4] = {0.01 * (k % 13 - 6), 0.5 + 0.01 * (k % 7), pl, std::sqrt(g * pl / (0.5 + 0.01 * (k % 7)))};
            const double r[4] = {-0.02 * (k % 11 - 5), 0.3 + 0.02 * (k % 5), pr, std::sqrt(g * pr / (0.3 + 0.02 * (k % 5)))};
            double ps = 0, vs = 0;
            orc_riemann(s, g, l, r, &ps, &vs);
            CHECK(std::isfinite(ps) && std::isfinite(vs), "riemann %d/%d", s, k);
        }
    // a negative-state input takes the solver's guard path
    const double bad[4] = {0.0, -1.0, 1.0, 1.0}, ok[4] = {0.0, 1.0, 1.0, 1.0};
    double ps = 0, vs = 0;
    orc_riemann(1, g, bad, ok, &ps, &vs);
    for (int d = 1; d <= 3; ++d)
        for (int kern = 0; kern < 3; ++kern) {
            if (d == 1 && kern == SPHGPU_WENDLAND) continue;
            double out[3];
            for (int q = 0; q <= 20; ++q) orc_kernel_eval(d, kern, 0.06 * q, 1.0, out);
        }
}

void run_io(const Case &c, const std::string &dir)
{
    const int n = c.n, d = c.prm.dim;
    const std::vector<int> rows = snapshot_rows(d);
    sphgpu_ctx *ctx = host_ctx_create(&c.prm, n, c.f.data());
    sphgpu_output_meta meta;
    CHECK(sphgpu_output_meta_init(ctx, 3, &meta) == 0, "%s: meta_init", c.name.c_str());
    meta.time_code = 0.125;
    char fname[256];
    CHECK(sphgpu_snapshot_filename(dir.c_str(), "snapshot", 3, ".csv", fname, sizeof fname) == 0, "filename");
    char tiny[8];
    CHECK(sphgpu_snapshot_filename(dir.c_str(), "snapshot", 3, ".csv", tiny, sizeof tiny) == SPHGPU_E_ARG,
          "filename into a short buffer");

    // CSV: host block and context paths, round trips, the oracle's writer
    const std::string csv = dir + "/" + c.name + ".csv", csv2 = dir + "/" + c.name + ".ctx.csv",
                      csvo = dir + "/" + c.name + ".orc.csv";
    CHECK(sphgpu_write_csv_fields(csv.c_str(), &meta, d, n, c.f.data()) == 0, "%s: write csv", c.name.c_str());
    CHECK(sphgpu_write_snapshot_csv(ctx, csv2.c_str(), &meta) == 0, "%s: write csv (ctx)", c.name.c_str());
    CHECK(orc_csv_write(csvo.c_str(), &meta, d, n, c.f.data()) == 0, "%s: oracle csv", c.name.c_str());
    CHECK(slurp(csv) == slurp(csv2), "%s: csv paths differ", c.name.c_str());
    CHECK(slurp(csv) == slurp(csvo), "%s: csv differs from the oracle's", c.name.c_str());
    int32_t nrows = 0;
    CHECK(sphgpu_csv_row_count(csv.c_str(), &nrows) == 0 && nrows == n, "%s: row count", c.name.c_str());
    std::vector<double> back((size_t)SPHGPU_NFIELDS * n, -1.0);
    sphgpu_output_meta mback;
    CHECK(sphgpu_read_csv_fields(csv.c_str(), &mback, d, n, back.data()) == 0, "%s: read csv", c.name.c_str());
    CHECK(rows_equal(back, c.f, n, rows), "%s: csv round trip", c.name.c_str());
    CHECK(mback.time_code == meta.time_code, "%s: csv time", c.name.c_str());
    std::vector<double> oback((size_t)SPHGPU_NFIELDS * n, 0.0);
    int orows = 0;
    CHECK(orc_csv_read(csv.c_str(), &mback, n, oback.data(), &orows) == 0 && orows == n, "%s: oracle csv read",
          c.name.c_str());
    CHECK(sphgpu_read_snapshot_csv(ctx, csv.c_str(), &mback) == 0, "%s: read csv (ctx)", c.name.c_str());
    CHECK(rows_equal(std::vector<double>(host_ctx_fields(ctx), host_ctx_fields(ctx) + back.size()), c.f, n, rows),
          "%s: csv ctx round trip", c.name.c_str());
    // inputs the reader rejects: short header, a short row, an unparsable
    // field, no particle, a missing file
    const std::string text = slurp(csv);
    const std::string bad = dir + "/bad.csv";
    size_t cut = 0;
    for (int k = 0; k < 20; ++k) cut = text.find('\n', cut) + 1;
    spit(bad, text.substr(0, cut));
    CHECK(sphgpu_read_csv_fields(bad.c_str(), nullptr, d, n, back.data()) == SPHGPU_E_ARG, "short header");
    const size_t last = text.rfind('\n', text.size() - 2);
    spit(bad, text.substr(0, last + 1) + "1,2,3\n");
    CHECK(sphgpu_read_csv_fields(bad.c_str(), nullptr, d, n, back.data()) == SPHGPU_E_ARG, "short row");
    std::string garbled = text;
    garbled[last + 3] = 'x';
    spit(bad, garbled);
    CHECK(sphgpu_read_csv_fields(bad.c_str(), nullptr, d, n, back.data()) == SPHGPU_E_ARG, "garbled field");
    const size_t head_end = text.find("\n", text.find("id,pos_x")) + 1;
    spit(bad, text.substr(0, head_end));
    CHECK(sphgpu_read_csv_fields(bad.c_str(), nullptr, d, n, back.data()) == SPHGPU_E_ARG, "no particle");
    CHECK(sphgpu_read_csv_fields((dir + "/absent.csv").c_str(), nullptr, d, n, back.data()) != 0, "missing file");

    // energy file
    const std::string en = dir + "/" + c.name + ".energy.dat";
    CHECK(sphgpu_energy_file_begin(en.c_str(), &meta) == 0, "energy begin");
    CHECK(sphgpu_energy_file_write(en.c_str(), 0.1, 1.0, 2.0, -3.0, 16) == 0, "energy write");
    CHECK(sphgpu_energy_file_append(ctx, en.c_str(), 16) == 0, "energy append");

    // VTK, binary and ASCII, host block and context; the empty set fails
    for (int binary = 0; binary < 2; ++binary) {
        const std::string v1 = dir + "/" + c.name + (binary ? ".b.vtk" : ".a.vtk");
        const std::string v2 = dir + "/" + c.name + (binary ? ".b.ctx.vtk" : ".a.ctx.vtk");
        CHECK(sphgpu_write_vtk_fields(v1.c_str(), &meta, d, n, c.f.data(), binary) == 0, "vtk");
        CHECK(sphgpu_write_snapshot_vtk(ctx, v2.c_str(), &meta, binary) == 0, "vtk ctx");
        CHECK(slurp(v1) == slurp(v2), "%s: vtk paths differ", c.name.c_str());
        CHECK(sphgpu_write_vtk_fields((dir + "/empty.vtk").c_str(), &meta, d, 0, c.f.data(), binary) == SPHGPU_E_ARG,
              "vtk empty");
    }

    // metadata JSON: size query, exact buffer, short buffer, round trip,
    // rejected texts
    for (int which = 0; which < 2; ++which) {
        int32_t need = 0;
        CHECK(sphgpu_output_meta_json(&meta, which, nullptr, 0, &need) == 0 && need > 1, "json size");
        std::vector<char> buf((size_t)need);
        CHECK(sphgpu_output_meta_json(&meta, which, buf.data(), need, &need) == 0, "json");
        CHECK(sphgpu_output_meta_json(&meta, which, buf.data(), need - 1, &need) != 0, "json short buffer");
        if (which == 0) {
            sphgpu_output_meta m2;
            CHECK(sphgpu_output_meta_from_json(buf.data(), &m2) == 0, "from_json");
            CHECK(m2.time_code == meta.time_code && m2.gamma == meta.gamma && m2.step == meta.step,
                  "json round trip");
        }
    }
    for (const char *t : {"", "{", "[1, 2]", "{\"step\": \"x\"}", "{\"a\": [1, {\"b\": nul}]}", "\"\\u12",
                          "{\"x\": 1} trailing"}) {
        sphgpu_output_meta m2;
        CHECK(sphgpu_output_meta_from_json(t, &m2) == SPHGPU_E_ARG, "from_json accepted '%s'", t);
    }

#ifdef WITH_HDF5
    for (int comp : {0, 4}) {
        const std::string h1 = dir + "/" + c.name + ".h5";
        CHECK(sphgpu_write_hdf5_fields(h1.c_str(), &meta, d, n, c.f.data(), comp) == 0, "hdf5 write");
        int32_t hr = 0;
        CHECK(sphgpu_hdf5_row_count(h1.c_str(), &hr) == 0 && hr == n, "hdf5 rows");
        std::vector<double> hb((size_t)SPHGPU_NFIELDS * n, -1.0);
        CHECK(sphgpu_read_hdf5_fields(h1.c_str(), &mback, n, hb.data()) == 0, "hdf5 read");
        CHECK(rows_equal(hb, c.f, n, rows), "%s: hdf5 round trip", c.name.c_str());
        CHECK(sphgpu_write_snapshot_hdf5(ctx, h1.c_str(), &meta, comp) == 0, "hdf5 write (ctx)");
        CHECK(sphgpu_read_snapshot_hdf5(ctx, h1.c_str(), &mback) == 0, "hdf5 read (ctx)");
        CHECK(sphgpu_write_hdf5_fields((dir + "/empty.h5").c_str(), &meta, d, 0, c.f.data(), comp) == SPHGPU_E_ARG,
              "hdf5 empty");
    }
    CHECK(sphgpu_read_hdf5_fields(csv.c_str(), nullptr, n, back.data()) != 0, "hdf5 reader on a CSV file");
#endif
    host_ctx_destroy(ctx);
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_check <dir> <case.bin>...\n");
        return 2;
    }
    const std::string dir = argv[1];
    run_solvers();
    for (int a = 2; a < argc; ++a) {
        Case c;
        if (!load_case(argv[a], c)) {
            std::printf("FAIL cannot read %s\n", argv[a]);
            return 1;
        }
        run_oracle(c);
        run_io(c, dir);
        std::printf("case %s (n=%d, dim=%d): done\n", c.name.c_str(), c.n, c.prm.dim);
    }
    std::printf(g_fail ? "host_check: %d failures\n" : "host_check: ok\n", g_fail);
    return g_fail ? 1 : 0;
}
