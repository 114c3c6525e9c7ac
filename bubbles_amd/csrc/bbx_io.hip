// bbx_io.hip -- frame I/O in the reference's particle serializer format
// (host code only).
//
// Writer: SaveSphParticleSet / PushParticleSetToFile
// (src/third/serializer.cpp:822-911): one block appended to the file,
//
//   FluidBegin
//   \t"Type" particles
//   \t"Count" <n>
//   \t"Format" <flags as letters p v d m b n l o>
//   \t"Spacing" <%g>
//   \tDataBegin
//   \t\t<fields of particle 0, %g, space separated>
//   ...
//   \tDataEnd
//   FluidEnd
//
// Reader: SerializerLoadParticles3 = SerializerFindFluidSection +
// _SerializerLoadParticles3 (src/third/serializer.cpp:130-190,444-560): the
// first FluidBegin block; "Count" and "Format" are read up to DataBegin,
// then one particle per line until DataEnd (empty and '#' lines skipped).
// Numbers go through the reference's ParseFloat
// (src/third/obj_loader.cpp:55-170,270-277), the digit-accumulating parser
// of tinyobjloader, restated in parse_double below.
//
// The formatting of large frames is split over host threads (each formats
// a contiguous range of particles with snprintf("%g")); the ranges are
// written in order, so the file is byte-identical to a serial write.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "bbx_common.h"

namespace bbx {
namespace {

// SerializerStringFromFlags (src/third/serializer.cpp:51-62)
std::string flags_string(int flags) {
    std::string s;
    if (flags & BBX_SER_POSITION) s += "p";
    if (flags & BBX_SER_VELOCITY) s += "v";
    if (flags & BBX_SER_DENSITY) s += "d";
    if (flags & BBX_SER_MASS) s += "m";
    if (flags & BBX_SER_BOUNDARY) s += "b";
    if (flags & BBX_SER_NORMAL) s += "n";
    if (flags & BBX_SER_LAYERS) s += "l";
    if (flags & BBX_SER_BOUNDARY_EXCLUSIVE) s += "o";
    return s;
}

// SerializerFlagsFromString (src/third/serializer.cpp:64-84); -1 on an
// unknown letter
int flags_from_string(const std::string &spec) {
    int f = 0;
    for (char ch : spec) {
        switch (ch) {
            case 'p': case 'P': f |= BBX_SER_POSITION; break;
            case 'v': case 'V': f |= BBX_SER_VELOCITY; break;
            case 'd': case 'D': f |= BBX_SER_DENSITY; break;
            case 'm': case 'M': f |= BBX_SER_MASS; break;
            case 'b': case 'B': f |= BBX_SER_BOUNDARY; break;
            case 'n': case 'N': f |= BBX_SER_NORMAL; break;
            case 'l': case 'L': f |= BBX_SER_LAYERS; break;
            case 'o': case 'O': f |= BBX_SER_BOUNDARY_EXCLUSIVE; break;
            case 'z': case 'Z': f |= BBX_SER_XYZ; break;
            default: return -1;
        }
    }
    return f;
}

void append_g(std::string &out, double v, bool space) {
    char buf[40];
    const int k = std::snprintf(buf, sizeof buf, space ? " %g" : "%g", v);
    out.append(buf, (size_t)k);
}

void append_d(std::string &out, int v, bool space) {
    char buf[24];
    const int k = std::snprintf(buf, sizeof buf, space ? " %d" : "%d", v);
    out.append(buf, (size_t)k);
}

struct Frame {
    int n;
    const double *pos, *vel, *dens;
    double mass;
    const int *boundary;
    int flags;
};

// PushParticleSetToFile's per-particle line (src/third/serializer.cpp:822-873)
void format_range(const Frame &f, int a, int b, std::string &out) {
    for (int i = a; i < b; i++) {
        if ((f.flags & BBX_SER_BOUNDARY_EXCLUSIVE) && f.boundary && !f.boundary[i]) continue;
        bool sp = false;
        out += "\t\t";
        if (f.flags & BBX_SER_POSITION) {
            for (int k = 0; k < 3; k++) append_g(out, f.pos[3 * (size_t)i + k], sp || k > 0);
            sp = true;
        }
        if (f.flags & BBX_SER_VELOCITY) {
            for (int k = 0; k < 3; k++) append_g(out, f.vel[3 * (size_t)i + k], sp || k > 0);
            sp = true;
        }
        if (f.flags & BBX_SER_DENSITY) {
            append_g(out, f.dens[i], sp);
            sp = true;
        }
        if (f.flags & BBX_SER_MASS) {
            append_g(out, f.mass, sp);
            sp = true;
        }
        if ((f.flags & BBX_SER_BOUNDARY) && f.boundary) {
            append_d(out, f.boundary[i], sp);
            sp = true;
        }
        out += "\n";
    }
}

// tinyobjloader's number parser as the reference's ParseFloat uses it
// (src/third/obj_loader.cpp:55-170): sign, integer digits accumulated as
// m = 10 m + d, fraction digits added as d * 10^-k (table for k < 8,
// pow beyond), optional exponent e, result m * 5^e * 2^e via ldexp.  No
// digit before the point, or an 'e' without digits, is a failure.
bool parse_double(const char *s, const char *end, double *out) {
#pragma clang fp contract(off)
    if (s >= end) return false;
    static const double kPow10[] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
    double m = 0.0;
    int e = 0, got = 0;
    char sign = '+', esign = '+';
    const char *c = s;
    if (*c == '+' || *c == '-') {
        sign = *c;
        c++;
    } else if (!(*c >= '0' && *c <= '9')) {
        return false;
    }
    while (c != end && *c >= '0' && *c <= '9') {
        m *= 10;
        m += (double)(*c - '0');
        c++;
        got++;
    }
    if (got == 0) return false;
    if (c != end) {
        if (*c == '.') {
            c++;
            int k = 1;
            while (c != end && *c >= '0' && *c <= '9') {
                m += (double)(*c - '0') * (k < 8 ? kPow10[k] : std::pow(10.0, -k));
                k++;
                c++;
            }
        } else if (!(*c == 'e' || *c == 'E')) {
            c = end;  // anything else ends the number
        }
        if (c != end && (*c == 'e' || *c == 'E')) {
            c++;
            if (c != end && (*c == '+' || *c == '-')) {
                esign = *c;
                c++;
            } else if (!(c != end && *c >= '0' && *c <= '9')) {
                return false;
            }
            int digits = 0;
            while (c != end && *c >= '0' && *c <= '9') {
                e = e * 10 + (*c - '0');
                c++;
                digits++;
            }
            if (esign == '-') e = -e;
            if (digits == 0) return false;
        }
    }
    const double mag = e ? std::ldexp(m * std::pow(5.0, e), e) : m;
    *out = sign == '+' ? mag : -mag;
    return true;
}

// ParseFloat: skip blanks, take the token up to the next blank, parse it
// (0 on failure), advance past it
double parse_float(const char **tok) {
    *tok += std::strspn(*tok, " \t");
    const char *end = *tok + std::strcspn(*tok, " \t\r");
    double v = 0.0;
    parse_double(*tok, end, &v);
    *tok = end;
    return v;
}

bool can_skip(char ch) { return ch == ' ' || ch == '\t'; }

void strip_eol(std::string &l) {
    if (!l.empty() && l.back() == '\n') l.pop_back();
    if (!l.empty() && l.back() == '\r') l.pop_back();
}

}  // namespace

int write_frame(const char *path, int flags, double spacing, int n, const double *pos, const double *vel,
                const double *dens, double mass, const int *boundary) {
    if (!path) return fail(BBX_E_INVALID, "null path");
    if (n < 0) return fail(BBX_E_INVALID, "negative particle count");
    if (flags & ~(BBX_SER_POSITION | BBX_SER_VELOCITY | BBX_SER_DENSITY | BBX_SER_MASS |
                  BBX_SER_BOUNDARY | BBX_SER_BOUNDARY_EXCLUSIVE))
        return fail(BBX_E_INVALID, "unsupported serializer flags (normals, layers, xyz)");
    if (((flags & BBX_SER_POSITION) && !pos) || ((flags & BBX_SER_VELOCITY) && !vel) ||
        ((flags & BBX_SER_DENSITY) && !dens))
        return fail(BBX_E_INVALID, "a requested field has no data");
    if ((flags & BBX_SER_BOUNDARY_EXCLUSIVE) && !boundary)
        return fail(BBX_E_INVALID, "boundary-exclusive output needs boundary flags");
    int count = n;
    if ((flags & BBX_SER_BOUNDARY_EXCLUSIVE) && boundary) {
        count = 0;
        for (int i = 0; i < n; i++) count += boundary[i] > 0 ? 1 : 0;
    }
    const Frame fr{n, pos, vel, dens, mass, boundary, flags};
    const int nthreads = n < 20000 ? 1 : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::string> parts((size_t)nthreads);
    {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; t++) {
            const int a = (int)((long long)n * t / nthreads), b = (int)((long long)n * (t + 1) / nthreads);
            pool.emplace_back([&, a, b, t] { format_range(fr, a, b, parts[(size_t)t]); });
        }
        for (auto &th : pool) th.join();
    }
    FILE *fp = std::fopen(path, "a+");
    if (!fp) return fail(BBX_E_INVALID, std::string("cannot open ") + path);
    std::fprintf(fp, "FluidBegin\n");
    std::fprintf(fp, "\t\"Type\" particles\n");
    std::fprintf(fp, "\t\"Count\" %d\n", count);
    std::fprintf(fp, "\t\"Format\" %s\n", flags_string(flags).c_str());
    std::fprintf(fp, "\t\"Spacing\" %g\n", spacing);
    std::fprintf(fp, "\tDataBegin\n");
    for (const std::string &p : parts) std::fwrite(p.data(), 1, p.size(), fp);
    std::fprintf(fp, "\tDataEnd\n");
    std::fprintf(fp, "FluidEnd\n");
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    if (!ok) return fail(BBX_E_INVALID, std::string("write failed: ") + path);
    return BBX_OK;
}

// One call scans the file: the header up to DataBegin (count, format), then
// the particle lines up to DataEnd.  With cap < 0 only the count and flags
// are returned.
int read_frame(const char *path, int cap, int *count, int *flags, double *pos, double *vel, double *dens,
               double *mass, int *boundary) {
    if (!path || !count || !flags) return fail(BBX_E_INVALID, "null argument");
    std::ifstream is(path);
    if (!is) return fail(BBX_E_INVALID, std::string("cannot open ") + path);
    // SerializerFindFluidSection / SerializerFluidProcessToken
    int in_region = 0, declared = 0;
    std::string format, line;
    bool data = false;
    while (!data && std::getline(is, line)) {
        line += '\n';
        std::string tok;
        auto process = [&](size_t i) {
            if (tok.empty()) return;
            if (!in_region) {
                if (tok == "FluidBegin") in_region = 1;
            } else if (tok == "\"Count\"") {
                std::string sub = line.substr(i + 1);
                size_t j = 0;
                while (j < sub.size() && can_skip(sub[j])) j++;
                const char *t = sub.c_str() + j;
                declared = (int)parse_float(&t);
            } else if (tok == "\"Format\"") {
                std::string sub = line.substr(i + 1);
                size_t j = 0;
                while (j < sub.size() && can_skip(sub[j])) j++;
                format.clear();
                while (j < sub.size() && !can_skip(sub[j]) && sub[j] != '\n' && sub[j] != '\r') format += sub[j++];
            } else if (tok == "DataBegin") {
                data = true;
            } else if (tok == "FluidEnd") {
                in_region = 0;
            }
            tok.clear();
        };
        for (size_t i = 0; i < line.size() && !data; i++) {
            const char ch = line[i];
            if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f') {
                process(i);
                continue;
            }
            tok += ch;
        }
    }
    if (!data) return fail(BBX_E_INVALID, "no FluidBegin ... DataBegin section");
    int fl = *flags;
    if (!format.empty()) {
        fl = flags_from_string(format);
        if (fl < 0) return fail(BBX_E_INVALID, "unknown letter in \"Format\" " + format);
    }
    *flags = fl;
    // _SerializerLoadParticles3
    int n = 0;
    bool end = false;
    while (std::getline(is, line)) {
        strip_eol(line);
        if (line.empty()) continue;
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') continue;
        if (std::string(t).find("DataEnd") != std::string::npos) {
            end = true;
            break;
        }
        double p[3] = {0, 0, 0}, v[3] = {0, 0, 0}, d = 0, m = 0;
        int b = 0;
        if (fl & BBX_SER_POSITION) {
            for (double &x : p) x = parse_float(&t);
            while (can_skip(t[0])) t++;
        }
        if (fl & BBX_SER_VELOCITY) {
            for (double &x : v) x = parse_float(&t);
            while (can_skip(t[0])) t++;
        }
        if (fl & BBX_SER_DENSITY) {
            d = parse_float(&t);
            while (can_skip(t[0])) t++;
        }
        if (fl & BBX_SER_MASS) {
            m = parse_float(&t);
            while (can_skip(t[0])) t++;
        }
        if (fl & BBX_SER_BOUNDARY) {
            b = (int)parse_float(&t);
            while (can_skip(t[0])) t++;
            if ((fl & BBX_SER_BOUNDARY_EXCLUSIVE) && !b) continue;
        }
        if (cap >= 0) {
            if (n >= cap) return fail(BBX_E_RANGE, "more particles in the file than the capacity given");
            if (pos) std::memcpy(pos + 3 * (size_t)n, p, sizeof p);
            if (vel) std::memcpy(vel + 3 * (size_t)n, v, sizeof v);
            if (dens) dens[n] = d;
            if (mass) mass[n] = m;
            if (boundary) boundary[n] = b;
        }
        n++;
    }
    if (!end) return fail(BBX_E_INVALID, "unterminated fluid description, missing DataEnd");
    (void)declared;  // the reference only reports a count mismatch
    *count = n;
    return BBX_OK;
}

}  // namespace bbx

extern "C" {

int bbx_serializer_write(const char *path, int flags, double spacing, int n, const double *positions,
                         const double *velocities, const double *densities, double mass,
                         const int *boundary) {
    return bbx::write_frame(path, flags, spacing, n, positions, velocities, densities, mass, boundary);
}

int bbx_serializer_read(const char *path, int capacity, int *count, int *flags, double *positions,
                        double *velocities, double *densities, double *masses, int *boundary) {
    return bbx::read_frame(path, capacity, count, flags, positions, velocities, densities, masses, boundary);
}

int bbx_serializer_parse_number(const char *text, double *value) {
    if (!text || !value) return bbx::fail(BBX_E_INVALID, "null argument");
    const char *t = text;
    *value = bbx::parse_float(&t);
    return BBX_OK;
}

}  // extern "C"
