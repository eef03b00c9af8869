// meta_json.cpp -- the snapshot metadata as JSON text, as the HDF5 writer
// stores it in its "metadata_json" and "unit_system" attributes, and back.
//
// Reference: OutputMetadata::to_json / from_json (src/output_metadata.cpp:
// 67-168) and UnitSystem::to_json / from_json (src/units.cpp:206-250),
// serialised by nlohmann::json's dump(2) (src/writers/hdf5_writer.cpp:
// 114-120): object keys in std::map order (byte order), two-space indent,
// integers in decimal, booleans as true/false, and doubles by nlohmann's
// to_chars: the Grisu2 digit generation of F. Loitsch, "Printing
// Floating-Point Numbers Quickly and Accurately with Integers" (PLDI 2010),
// with the interval bounds shrunk by one unit and the last digit rounded
// towards the value ("round_weed" without the exactness check), laid out as
// fixed notation for decimal exponents in (-4, 15] and as d.ddde+XX
// otherwise, ".0" appended to integral values.  Grisu2 does not always give
// the shortest or closest digit string, so %.17g or a shortest round-trip
// search would print different text for about 1 value in 1600.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sphgpu.h"

namespace {

// ---- Grisu2 -----------------------------------------------------------------

struct Fp {
    uint64_t f;
    int e;
};

Fp fp_sub(Fp x, Fp y) { return {x.f - y.f, x.e}; }

// upper 64 bits of the 128-bit product, rounded half up
Fp fp_mul(Fp x, Fp y)
{
    const unsigned __int128 p = (unsigned __int128)x.f * y.f + ((unsigned __int128)1 << 63);
    return {(uint64_t)(p >> 64), x.e + y.e + 64};
}

Fp fp_normalize(Fp x)
{
    while (!(x.f >> 63)) {
        x.f <<= 1;
        --x.e;
    }
    return x;
}

// 10^k for k = -300, -292, ..., 324: the significand rounded to nearest into
// [2^63, 2^64) and its binary exponent
struct Pow10 {
    uint64_t f;
    int e, k;
};
constexpr Pow10 kPow10[] = {
    {0xAB70FE17C79AC6CAull, -1060, -300}, {0xFF77B1FCBEBCDC4Full, -1034, -292},
    {0xBE5691EF416BD60Cull, -1007, -284}, {0x8DD01FAD907FFC3Cull, -980, -276},
    {0xD3515C2831559A83ull, -954, -268}, {0x9D71AC8FADA6C9B5ull, -927, -260},
    {0xEA9C227723EE8BCBull, -901, -252}, {0xAECC49914078536Dull, -874, -244},
    {0x823C12795DB6CE57ull, -847, -236}, {0xC21094364DFB5637ull, -821, -228},
    {0x9096EA6F3848984Full, -794, -220}, {0xD77485CB25823AC7ull, -768, -212},
    {0xA086CFCD97BF97F4ull, -741, -204}, {0xEF340A98172AACE5ull, -715, -196},
    {0xB23867FB2A35B28Eull, -688, -188}, {0x84C8D4DFD2C63F3Bull, -661, -180},
    {0xC5DD44271AD3CDBAull, -635, -172}, {0x936B9FCEBB25C996ull, -608, -164},
    {0xDBAC6C247D62A584ull, -582, -156}, {0xA3AB66580D5FDAF6ull, -555, -148},
    {0xF3E2F893DEC3F126ull, -529, -140}, {0xB5B5ADA8AAFF80B8ull, -502, -132},
    {0x87625F056C7C4A8Bull, -475, -124}, {0xC9BCFF6034C13053ull, -449, -116},
    {0x964E858C91BA2655ull, -422, -108}, {0xDFF9772470297EBDull, -396, -100},
    {0xA6DFBD9FB8E5B88Full, -369, -92}, {0xF8A95FCF88747D94ull, -343, -84},
    {0xB94470938FA89BCFull, -316, -76}, {0x8A08F0F8BF0F156Bull, -289, -68},
    {0xCDB02555653131B6ull, -263, -60}, {0x993FE2C6D07B7FACull, -236, -52},
    {0xE45C10C42A2B3B06ull, -210, -44}, {0xAA242499697392D3ull, -183, -36},
    {0xFD87B5F28300CA0Eull, -157, -28}, {0xBCE5086492111AEBull, -130, -20},
    {0x8CBCCC096F5088CCull, -103, -12}, {0xD1B71758E219652Cull, -77, -4},
    {0x9C40000000000000ull, -50, 4}, {0xE8D4A51000000000ull, -24, 12},
    {0xAD78EBC5AC620000ull, 3, 20}, {0x813F3978F8940984ull, 30, 28},
    {0xC097CE7BC90715B3ull, 56, 36}, {0x8F7E32CE7BEA5C70ull, 83, 44},
    {0xD5D238A4ABE98068ull, 109, 52}, {0x9F4F2726179A2245ull, 136, 60},
    {0xED63A231D4C4FB27ull, 162, 68}, {0xB0DE65388CC8ADA8ull, 189, 76},
    {0x83C7088E1AAB65DBull, 216, 84}, {0xC45D1DF942711D9Aull, 242, 92},
    {0x924D692CA61BE758ull, 269, 100}, {0xDA01EE641A708DEAull, 295, 108},
    {0xA26DA3999AEF774Aull, 322, 116}, {0xF209787BB47D6B85ull, 348, 124},
    {0xB454E4A179DD1877ull, 375, 132}, {0x865B86925B9BC5C2ull, 402, 140},
    {0xC83553C5C8965D3Dull, 428, 148}, {0x952AB45CFA97A0B3ull, 455, 156},
    {0xDE469FBD99A05FE3ull, 481, 164}, {0xA59BC234DB398C25ull, 508, 172},
    {0xF6C69A72A3989F5Cull, 534, 180}, {0xB7DCBF5354E9BECEull, 561, 188},
    {0x88FCF317F22241E2ull, 588, 196}, {0xCC20CE9BD35C78A5ull, 614, 204},
    {0x98165AF37B2153DFull, 641, 212}, {0xE2A0B5DC971F303Aull, 667, 220},
    {0xA8D9D1535CE3B396ull, 694, 228}, {0xFB9B7CD9A4A7443Cull, 720, 236},
    {0xBB764C4CA7A44410ull, 747, 244}, {0x8BAB8EEFB6409C1Aull, 774, 252},
    {0xD01FEF10A657842Cull, 800, 260}, {0x9B10A4E5E9913129ull, 827, 268},
    {0xE7109BFBA19C0C9Dull, 853, 276}, {0xAC2820D9623BF429ull, 880, 284},
    {0x80444B5E7AA7CF85ull, 907, 292}, {0xBF21E44003ACDD2Dull, 933, 300},
    {0x8E679C2F5E44FF8Full, 960, 308}, {0xD433179D9C8CB841ull, 986, 316},
    {0x9E19DB92B4E31BA9ull, 1013, 324},
};
constexpr int kAlpha = -60, kGamma = -32;

// a cached power c = 10^-k with alpha <= e + c.e + 64 <= gamma
Pow10 cached_power(int e)
{
    const int f = kAlpha - e - 1;
    // ceil(f log10(2)) for the exponent range of doubles
    const int k = (f * 78913) / (1 << 18) + (f > 0 ? 1 : 0);
    const int idx = (300 + k + 7) / 8;
    return kPow10[idx];
}

// the largest power of ten <= n (n > 0) and its digit count
int pow10_below(uint32_t n, uint32_t &p)
{
    static const uint32_t t[] = {1u, 10u, 100u, 1000u, 10000u, 100000u, 1000000u, 10000000u, 100000000u,
                                 1000000000u};
    int d = 10;
    while (d > 1 && n < t[d - 1]) --d;
    p = t[d - 1];
    return d;
}

void round_last(char *buf, int len, uint64_t dist, uint64_t delta, uint64_t rest, uint64_t ten_k)
{
    // move the last digit towards w while the candidate stays inside the
    // interval and gets closer to w
    while (rest < dist && delta - rest >= ten_k && (rest + ten_k < dist || dist - rest > rest + ten_k - dist)) {
        --buf[len - 1];
        rest += ten_k;
    }
}

// digits of a value in (M-, M+) closest to w: buf gets len digits, the value
// is digits * 10^dexp
void digit_gen(char *buf, int &len, int &dexp, Fp Mm, Fp w, Fp Mp)
{
    uint64_t delta = fp_sub(Mp, Mm).f;
    uint64_t dist = fp_sub(Mp, w).f;
    const int sh = -Mp.e;
    const uint64_t one = 1ull << sh;
    uint32_t p1 = (uint32_t)(Mp.f >> sh);
    uint64_t p2 = Mp.f & (one - 1);
    uint32_t pw;
    int n = pow10_below(p1, pw);
    while (n > 0) {
        const uint32_t d = p1 / pw;
        p1 %= pw;
        buf[len++] = (char)('0' + d);
        --n;
        const uint64_t rest = ((uint64_t)p1 << sh) + p2;
        if (rest <= delta) {
            dexp += n;
            round_last(buf, len, dist, delta, rest, (uint64_t)pw << sh);
            return;
        }
        pw /= 10;
    }
    int m = 0;
    for (;;) {
        p2 *= 10;
        delta *= 10;
        dist *= 10;
        buf[len++] = (char)('0' + (p2 >> sh));
        p2 &= one - 1;
        ++m;
        if (p2 <= delta) break;
    }
    dexp -= m;
    round_last(buf, len, dist, delta, p2, one);
}

// v > 0, finite
void grisu2(char *buf, int &len, int &dexp, double v)
{
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    const uint64_t F = bits & ((1ull << 52) - 1);
    const int E = (int)(bits >> 52);
    const Fp x = E == 0 ? Fp{F, 1 - 1075} : Fp{F | (1ull << 52), E - 1075};
    const bool lower_closer = F == 0 && E > 1;
    const Fp mp = fp_normalize(Fp{2 * x.f + 1, x.e - 1});
    Fp mm = lower_closer ? Fp{4 * x.f - 1, x.e - 2} : Fp{2 * x.f - 1, x.e - 1};
    mm.f <<= (mm.e - mp.e);
    mm.e = mp.e;
    const Fp w0 = fp_normalize(x);
    const Pow10 c = cached_power(mp.e);
    const Fp cp{c.f, c.e};
    const Fp w = fp_mul(w0, cp), wm = fp_mul(mm, cp), wp = fp_mul(mp, cp);
    len = 0;
    dexp = -c.k;
    digit_gen(buf, len, dexp, Fp{wm.f + 1, wm.e}, w, Fp{wp.f - 1, wp.e});
}

std::string json_double(double v)
{
    if (!std::isfinite(v)) return "null";
    std::string s;
    if (std::signbit(v)) {
        s += '-';
        v = -v;
    }
    if (v == 0.0) return s + "0.0";
    char d[32];
    int k = 0, dexp = 0;
    grisu2(d, k, dexp, v);
    const int n = k + dexp;  // the decimal point sits after digit n
    if (k <= n && n <= 15) {
        s.append(d, (size_t)k);
        s.append((size_t)(n - k), '0');
        s += ".0";
    } else if (0 < n && n <= 15) {
        s.append(d, (size_t)n);
        s += '.';
        s.append(d + n, (size_t)(k - n));
    } else if (-4 < n && n <= 0) {
        s += "0.";
        s.append((size_t)(-n), '0');
        s.append(d, (size_t)k);
    } else {
        s += d[0];
        if (k > 1) {
            s += '.';
            s.append(d + 1, (size_t)(k - 1));
        }
        int e = n - 1;
        s += 'e';
        s += e < 0 ? '-' : '+';
        e = std::abs(e);
        char eb[8];
        std::snprintf(eb, sizeof(eb), e < 10 ? "0%d" : "%d", e);
        s += eb;
    }
    return s;
}

std::string json_string(const char *t)
{
    std::string s = "\"";
    for (const unsigned char *p = (const unsigned char *)t; *p; ++p) {
        switch (*p) {
        case '"': s += "\\\""; break;
        case '\\': s += "\\\\"; break;
        case '\b': s += "\\b"; break;
        case '\f': s += "\\f"; break;
        case '\n': s += "\\n"; break;
        case '\r': s += "\\r"; break;
        case '\t': s += "\\t"; break;
        default:
            if (*p < 0x20) {
                char u[8];
                std::snprintf(u, sizeof(u), "\\u%04x", *p);
                s += u;
            } else {
                s += (char)*p;
            }
        }
    }
    return s + "\"";
}

// an object as nlohmann's dump(2) prints it: keys in byte order
struct Obj {
    std::map<std::string, std::string> v;  // key -> rendered value (objects pre-rendered)
    void num(const char *k, double x) { v[k] = json_double(x); }
    void integer(const char *k, long long x) { v[k] = std::to_string(x); }
    void boolean(const char *k, bool x) { v[k] = x ? "true" : "false"; }
    void str(const char *k, const char *x) { v[k] = json_string(x); }
    std::string dump(int level) const
    {
        if (v.empty()) return "{}";
        std::string s = "{\n";
        size_t i = 0;
        for (const auto &kv : v) {
            s.append((size_t)(2 * (level + 1)), ' ');
            s += json_string(kv.first.c_str());
            s += ": ";
            s += kv.second;
            s += (++i < v.size()) ? ",\n" : "\n";
        }
        s.append((size_t)(2 * level), ' ');
        return s + "}";
    }
};

int unit_type_enum(const char *name)
{
    if (!std::strcmp(name, "Code")) return 0;
    if (!std::strcmp(name, "Galactic")) return 1;
    if (!std::strcmp(name, "SI")) return 2;
    if (!std::strcmp(name, "CGS")) return 3;
    return -1;
}

// UnitSystem's names per type (src/units.cpp:125-203)
const char *const kUnitNames[5][8] = {
    {"Code", "code_length", "code_mass", "code_time", "code_velocity", "code_energy", "code_density",
     "code_pressure"},
    {"Galactic", "kpc", "M_sun", "Myr", "km/s", "erg", "M_sun/kpc^3", "erg/cm^3"},
    {"SI", "m", "kg", "s", "m/s", "J", "kg/m^3", "Pa"},
    {"CGS", "cm", "g", "s", "cm/s", "erg", "g/cm^3", "dyne/cm^2"},
    {"Unknown", "unknown", "unknown", "unknown", "unknown", "unknown", "unknown", "unknown"},
};

// UnitSystem::to_json (src/units.cpp:206-231); the energy, density and
// pressure factors follow compute_derived_factors (src/units.cpp:55-60)
Obj units_obj(const sphgpu_output_meta &m)
{
    const int t = unit_type_enum(m.unit_type);
    const char *const *nm = kUnitNames[t < 0 ? 4 : t];
    Obj u;
    u.str("type", m.unit_type);
    u.integer("type_enum", t < 0 ? 4 : t);
    u.str("length_unit", m.length_unit);
    u.str("mass_unit", m.mass_unit);
    u.str("time_unit", m.time_unit);
    u.str("velocity_unit", m.velocity_unit);
    u.str("energy_unit", m.energy_unit);
    u.str("density_unit", nm[6]);
    u.str("pressure_unit", nm[7]);
    const double L = m.length_to_cgs, M = m.mass_to_cgs, V = m.velocity_to_cgs;
    const double E = M * V * V;
    u.num("length_to_cgs", L);
    u.num("mass_to_cgs", M);
    u.num("time_to_cgs", m.time_to_cgs);
    u.num("velocity_to_cgs", V);
    u.num("energy_to_cgs", E);
    u.num("density_to_cgs", M / (L * L * L));
    u.num("pressure_to_cgs", E / (L * L * L));
    return u;
}

const char *sph_name(int t)
{
    switch (t) {
    case SPHGPU_SSPH: return "Standard SPH";
    case SPHGPU_DISPH: return "Density Independent SPH";
    case SPHGPU_GSPH: return "Godunov SPH";
    default: return "Unknown";
    }
}

const char *kernel_type_name(int k)
{
    switch (k) {
    case SPHGPU_CUBIC_SPLINE: return "Cubic Spline";
    case SPHGPU_WENDLAND: return "Wendland C4";
    default: return "Unknown";
    }
}

// OutputMetadata::to_json (src/output_metadata.cpp:67-116)
std::string meta_json(const sphgpu_output_meta &m)
{
    Obj j;
    j.str("simulation_name", m.simulation_name);
    j.str("timestamp", m.timestamp);
    j.integer("format_version", m.format_version);
    j.integer("step", m.step);
    j.num("time_code", m.time_code);
    j.num("time_physical", m.time_physical);
    j.integer("particle_count", m.particle_count);
    j.str("sph_type", sph_name(m.sph_type));
    j.integer("sph_type_enum", m.sph_type);
    j.str("kernel_type", kernel_type_name(m.kernel));
    j.integer("kernel_type_enum", m.kernel);
    j.num("gamma", m.gamma);
    j.num("gravitational_constant", m.gravitational_constant);
    j.integer("neighbor_number", m.neighbor_number);
    j.boolean("use_gravity", m.use_gravity != 0);
    j.boolean("use_balsara", m.use_balsara != 0);
    j.boolean("use_time_dependent_av", m.use_time_dependent_av != 0);
    j.v["units"] = units_obj(m).dump(1);
    j.boolean("is_relaxation", m.is_relaxation != 0);
    if (m.is_relaxation) {
        j.integer("relaxation_step", m.relaxation_step);
        j.integer("relaxation_total_steps", m.relaxation_total_steps);
        j.num("accumulated_time", m.accumulated_time);
        j.num("alpha_scaling", m.alpha_scaling);
        j.num("rho_center", m.rho_center);
        j.num("K", m.K);
        j.num("R", m.R);
        j.num("M_total", m.M_total);
    }
    j.num("kinetic_energy", m.kinetic_energy);
    j.num("thermal_energy", m.thermal_energy);
    j.num("potential_energy", m.potential_energy);
    j.num("total_energy", m.total_energy);
    return j.dump(0);
}

// ---- a small JSON reader (objects, strings, numbers, literals) --------------

struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, OBJ, ARR } kind = NUL;
    bool b = false;
    bool is_int = false;
    double num = 0.0;
    long long inum = 0;
    std::string str;
    std::map<std::string, JVal> obj;
    std::vector<JVal> arr;
};

struct Parser {
    const char *p, *e;
    void ws()
    {
        while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool lit(const char *w)
    {
        const size_t k = std::strlen(w);
        if ((size_t)(e - p) < k || std::strncmp(p, w, k)) return false;
        p += k;
        return true;
    }
    bool string(std::string &out)
    {
        if (p >= e || *p != '"') return false;
        ++p;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) return false;
                switch (*p) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    if (e - p < 5) return false;
                    char h[5] = {p[1], p[2], p[3], p[4], 0};
                    const long c = std::strtol(h, nullptr, 16);
                    if (c >= 0x80) return false;  // ASCII escapes only
                    out += (char)c;
                    p += 4;
                    break;
                }
                default: return false;
                }
                ++p;
            } else {
                out += *p++;
            }
        }
        if (p >= e) return false;
        ++p;
        return true;
    }
    bool value(JVal &v)
    {
        ws();
        if (p >= e) return false;
        if (*p == '{') {
            ++p;
            v.kind = JVal::OBJ;
            ws();
            if (p < e && *p == '}') {
                ++p;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p >= e || *p != ':') return false;
                ++p;
                if (!value(v.obj[k])) return false;
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == '}') {
                    ++p;
                    return true;
                }
                return false;
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = JVal::ARR;
            ws();
            if (p < e && *p == ']') {
                ++p;
                return true;
            }
            for (;;) {
                v.arr.emplace_back();
                if (!value(v.arr.back())) return false;
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == ']') {
                    ++p;
                    return true;
                }
                return false;
            }
        }
        if (*p == '"') {
            v.kind = JVal::STR;
            return string(v.str);
        }
        if (lit("true")) { v.kind = JVal::BOOL; v.b = true; return true; }
        if (lit("false")) { v.kind = JVal::BOOL; v.b = false; return true; }
        if (lit("null")) { v.kind = JVal::NUL; return true; }
        const char *s = p;
        if (p < e && *p == '-') ++p;
        bool frac = false;
        while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) {
            if (*p == '.' || *p == 'e' || *p == 'E') frac = true;
            ++p;
        }
        if (p == s) return false;
        const std::string t(s, p);
        char *end = nullptr;
        v.kind = JVal::NUM;
        v.num = std::strtod(t.c_str(), &end);
        if (end != t.c_str() + t.size()) return false;
        v.is_int = !frac;
        if (v.is_int) v.inum = std::strtoll(t.c_str(), nullptr, 10);
        return true;
    }
};

const JVal *find(const JVal &o, const char *k)
{
    if (o.kind != JVal::OBJ) return nullptr;
    auto it = o.obj.find(k);
    return it == o.obj.end() ? nullptr : &it->second;
}

// j.value(key, default) for the types from_json asks for; a key of another
// type makes nlohmann throw (type_error 302), reported as false
struct Getter {
    const JVal &o;
    bool ok = true;
    void d(const char *k, double &out, double def)
    {
        const JVal *v = find(o, k);
        if (!v) { out = def; return; }
        if (v->kind != JVal::NUM) { ok = false; return; }
        out = v->num;
    }
    void i(const char *k, int32_t &out, int32_t def)
    {
        const JVal *v = find(o, k);
        if (!v) { out = def; return; }
        if (v->kind != JVal::NUM) { ok = false; return; }
        out = v->is_int ? (int32_t)v->inum : (int32_t)v->num;
    }
    void b(const char *k, int32_t &out, bool def)
    {
        const JVal *v = find(o, k);
        if (!v) { out = def; return; }
        if (v->kind != JVal::BOOL) { ok = false; return; }
        out = v->b;
    }
    void s(const char *k, char *out, size_t cap, const char *def)
    {
        const JVal *v = find(o, k);
        const std::string t = !v ? std::string(def) : v->str;
        if (v && v->kind != JVal::STR) { ok = false; return; }
        const size_t n = std::min(cap - 1, t.size());
        std::memcpy(out, t.data(), n);
        out[n] = '\0';
    }
};

} // namespace

extern "C" {

int sphgpu_output_meta_json(const sphgpu_output_meta *meta, int32_t which, char *buf, int32_t buf_len,
                            int32_t *needed)
{
    if (!meta || (which != 0 && which != 1) || buf_len < 0 || (buf_len > 0 && !buf)) return SPHGPU_E_ARG;
    const std::string s = which == 0 ? meta_json(*meta) : units_obj(*meta).dump(0);
    if (needed) *needed = (int32_t)s.size() + 1;
    if ((int32_t)s.size() + 1 > buf_len) return buf_len == 0 ? 0 : SPHGPU_E_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

int sphgpu_output_meta_from_json(const char *text, sphgpu_output_meta *m)
{
    if (!text || !m) return SPHGPU_E_ARG;
    JVal j;
    Parser ps{text, text + std::strlen(text)};
    if (!ps.value(j) || j.kind != JVal::OBJ) return SPHGPU_E_ARG;
    ps.ws();
    if (ps.p != ps.e) return SPHGPU_E_ARG;
    // OutputMetadata::from_json (src/output_metadata.cpp:118-168) starts from
    // OutputMetadata(): the defaults of every absent key
    sphgpu_output_meta d;
    std::memset(&d, 0, sizeof(d));
    Getter g{j};
    // an absent timestamp defaults to the current UTC time, as
    // OutputMetadata::generate_timestamp (src/output_metadata.cpp:41-48)
    char ts[32];
    const std::time_t now = std::time(nullptr);
    std::tm tm_utc;
    gmtime_r(&now, &tm_utc);
    std::strftime(ts, sizeof(ts), "%Y-%m-%dT%H:%M:%SZ", &tm_utc);
    g.s("simulation_name", d.simulation_name, sizeof(d.simulation_name), "SPH Simulation");
    g.s("timestamp", d.timestamp, sizeof(d.timestamp), ts);
    g.i("format_version", d.format_version, 1);
    g.i("step", d.step, 0);
    g.d("time_code", d.time_code, 0.0);
    g.d("time_physical", d.time_physical, 0.0);
    g.i("particle_count", d.particle_count, 0);
    g.i("sph_type_enum", d.sph_type, 0);
    g.i("kernel_type_enum", d.kernel, 0);
    g.d("gamma", d.gamma, 1.4);
    g.d("gravitational_constant", d.gravitational_constant, 1.0);
    g.i("neighbor_number", d.neighbor_number, 32);
    g.b("use_gravity", d.use_gravity, false);
    g.b("use_balsara", d.use_balsara, true);
    g.b("use_time_dependent_av", d.use_time_dependent_av, false);
    // UnitSystem::from_json (src/units.cpp:233-250): the type and the factors;
    // the names follow the type
    int t = 0;
    d.length_to_cgs = d.mass_to_cgs = d.time_to_cgs = d.velocity_to_cgs = 1.0;
    if (const JVal *u = find(j, "units")) {
        Getter gu{*u};
        int32_t te = 0;
        gu.i("type_enum", te, 0);
        t = te;
        gu.d("length_to_cgs", d.length_to_cgs, 1.0);
        gu.d("mass_to_cgs", d.mass_to_cgs, 1.0);
        gu.d("time_to_cgs", d.time_to_cgs, 1.0);
        gu.d("velocity_to_cgs", d.velocity_to_cgs, 1.0);
        if (!gu.ok || u->kind != JVal::OBJ) return SPHGPU_E_ARG;
    }
    const char *const *nm = kUnitNames[(t < 0 || t > 3) ? 4 : t];
    std::snprintf(d.unit_type, sizeof(d.unit_type), "%s", nm[0]);
    std::snprintf(d.length_unit, sizeof(d.length_unit), "%s", nm[1]);
    std::snprintf(d.mass_unit, sizeof(d.mass_unit), "%s", nm[2]);
    std::snprintf(d.time_unit, sizeof(d.time_unit), "%s", nm[3]);
    std::snprintf(d.velocity_unit, sizeof(d.velocity_unit), "%s", nm[4]);
    std::snprintf(d.energy_unit, sizeof(d.energy_unit), "%s", nm[5]);
    g.b("is_relaxation", d.is_relaxation, false);
    d.alpha_scaling = d.rho_center = d.K = d.R = d.M_total = 1.0;
    if (d.is_relaxation) {
        g.i("relaxation_step", d.relaxation_step, 0);
        g.i("relaxation_total_steps", d.relaxation_total_steps, 0);
        g.d("accumulated_time", d.accumulated_time, 0.0);
        g.d("alpha_scaling", d.alpha_scaling, 1.0);
        g.d("rho_center", d.rho_center, 0.0);
        g.d("K", d.K, 0.0);
        g.d("R", d.R, 0.0);
        g.d("M_total", d.M_total, 0.0);
    }
    g.d("kinetic_energy", d.kinetic_energy, 0.0);
    g.d("thermal_energy", d.thermal_energy, 0.0);
    g.d("potential_energy", d.potential_energy, 0.0);
    g.d("total_energy", d.total_energy, 0.0);
    if (!g.ok) return SPHGPU_E_ARG;
    d.precision = 16;
    *m = d;
    return 0;
}

} // extern "C"
