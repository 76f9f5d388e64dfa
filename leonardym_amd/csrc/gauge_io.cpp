// gauge_io.cpp -- readers and writers of the reference's gauge-configuration
// files, on host arrays of fundamental gauge links (gaugeLinkConfiguration raw
// data: V x [4 mu][N x N column-major] complex double, sites t-fastest).
//
// Formats (all numbers XDR: 4-byte big-endian ints, big-endian IEEE floats
// and doubles, RFC 4506):
//   leonard_format  (ReadStartGaugeConfiguration.cpp:51-130, OutputSweep.cpp:32-84)
//     <dir><name>_<n>.descriptor.txt : "Lx Ly Lz Lt\n px py pz pt\n nproc\n localsize\n plaquette"
//     <dir><name>_<n>_<proc>.txt     : localsize records of
//         int x, y, z, t; then for mu, i (row), j (column): double re, im
//   muenster_format, SU(2) (ReadStartGaugeConfiguration.cpp:132-300, OutputSweep.cpp:87-283)
//     <dir>0:<n>:<name> : sites with x fastest, then y, z, t; per mu the first
//         row of U^+ as two (re, im) float pairs; then a trailer: plaquette,
//         0, 0, beta (floats); kappas (all but the last negated) with alphas
//         and (epsilon, lambda) pairs (floats); the sweep, 5 polynomial orders
//         per kappa, Lx Ly Lz Lt and lvrc (ints)
//   muenster_format, SU(3) (ReadStartGaugeConfiguration.cpp:300-480, OutputSweep.cpp:284-440)
//     <dir><name>.config-<n> : the same site order; per mu the first two
//         columns of U^* as (re, -im) float pairs, rows outer; on reading
//         the two columns are re-orthonormalised (Gram-Schmidt) and the
//         third is conj(col0 x col1); trailer Lx Ly Lz Lt (ints), beta, c1
//         and the temporal plaquette (doubles)
//
// Deliberate differences from the reference: a short or unreadable file is an
// error (the reference ignores the XDR return codes of the leonard reader and
// keeps whatever it had), and nothing is printed.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lym.h"

namespace lym_internal {
int fail(int code, const char* fmt, ...);
}
using lym_internal::fail;

namespace {

typedef std::complex<double> cd;

// ---------------------------------------------------------------- XDR streams
struct XdrOut {
    std::vector<unsigned char> buf;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) buf.push_back((unsigned char)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) buf.push_back((unsigned char)(v >> s));
    }
    void i32(int v) { u32((uint32_t)v); }
    void f32(float f) {
        uint32_t v;
        std::memcpy(&v, &f, 4);
        u32(v);
    }
    void f64(double d) {
        uint64_t v;
        std::memcpy(&v, &d, 8);
        u64(v);
    }
};

struct XdrIn {
    const unsigned char* p;
    size_t n, pos;
    bool ok = true;
    uint32_t u32() {
        if (pos + 4 > n) {
            ok = false;
            return 0;
        }
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) v = (v << 8) | p[pos + k];
        pos += 4;
        return v;
    }
    uint64_t u64() {
        if (pos + 8 > n) {
            ok = false;
            return 0;
        }
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k) v = (v << 8) | p[pos + k];
        pos += 8;
        return v;
    }
    int i32() { return (int)u32(); }
    float f32() {
        uint32_t v = u32();
        float f;
        std::memcpy(&f, &v, 4);
        return f;
    }
    double f64() {
        uint64_t v = u64();
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
};

int read_file(const std::string& path, std::vector<unsigned char>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(LYM_ERR_INVALID, "cannot open %s for reading", path.c_str());
    out.clear();
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) out.insert(out.end(), chunk, chunk + got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return fail(LYM_ERR_INVALID, "read error on %s", path.c_str());
    return LYM_OK;
}

int write_file(const std::string& path, const void* data, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(LYM_ERR_INVALID, "cannot open %s for writing", path.c_str());
    const size_t put = n ? std::fwrite(data, 1, n, f) : 0;
    const int closed = std::fclose(f);
    if (put != n || closed != 0) return fail(LYM_ERR_INVALID, "write error on %s", path.c_str());
    return LYM_OK;
}

struct Lat {
    int L[4];
    long V;
    int N;
    long lex(int x, int y, int z, int t) const { return ((((long)x * L[1] + y) * L[2] + z) * L[3]) + t; }
    // element (i, j) of U_mu(site): column-major
    size_t at(long site, int mu, int i, int j) const { return ((size_t)(site * 4 + mu) * N * N + (size_t)j * N + i) * 2; }
};

int make_lat(const int dims[4], int ncolors, Lat& l) {
    if (!dims) return fail(LYM_ERR_INVALID, "null dims");
    if (ncolors != 2 && ncolors != 3) return fail(LYM_ERR_INVALID, "ncolors must be 2 or 3 (got %d)", ncolors);
    l.V = 1;
    for (int m = 0; m < 4; ++m) {
        if (dims[m] < 1) return fail(LYM_ERR_INVALID, "lattice extent %d", dims[m]);
        l.L[m] = dims[m];
        l.V *= dims[m];
    }
    l.N = ncolors;
    return LYM_OK;
}

cd link_el(const Lat& l, const double* U, long s, int mu, int i, int j) {
    const size_t k = l.at(s, mu, i, j);
    return cd(U[k], U[k + 1]);
}
void set_el(const Lat& l, double* U, long s, int mu, int i, int j, cd v) {
    const size_t k = l.at(s, mu, i, j);
    U[k] = v.real();
    U[k + 1] = v.imag();
}

// Re tr of the (mu, t) plaquettes summed over sites and mu < 3, divided by
// 3 N V (Plaquette::temporalPlaquette, source/wilson_loops/Plaquette.cpp:49-66)
double temporal_plaquette(const Lat& l, const double* U) {
    const int N = l.N;
    long double sum = 0;
    for (int x = 0; x < l.L[0]; ++x)
        for (int y = 0; y < l.L[1]; ++y)
            for (int z = 0; z < l.L[2]; ++z)
                for (int t = 0; t < l.L[3]; ++t) {
                    const int c[4] = {x, y, z, t};
                    const long s = l.lex(x, y, z, t);
                    int ct[4] = {x, y, z, (t + 1) % l.L[3]};
                    const long s_t = l.lex(ct[0], ct[1], ct[2], ct[3]);
                    for (int mu = 0; mu < 3; ++mu) {
                        int cm[4] = {c[0], c[1], c[2], c[3]};
                        cm[mu] = (cm[mu] + 1) % l.L[mu];
                        const long s_mu = l.lex(cm[0], cm[1], cm[2], cm[3]);
                        // tr[U_mu(s) U_t(s+mu) U_mu(s+t)^+ U_t(s)^+]
                        cd a[3][3], b[3][3];
                        for (int i = 0; i < N; ++i)
                            for (int j = 0; j < N; ++j) {
                                cd v = 0;
                                for (int k = 0; k < N; ++k) v += link_el(l, U, s, mu, i, k) * link_el(l, U, s_mu, 3, k, j);
                                a[i][j] = v;
                            }
                        for (int i = 0; i < N; ++i)
                            for (int j = 0; j < N; ++j) {
                                cd v = 0;
                                for (int k = 0; k < N; ++k)
                                    v += std::conj(link_el(l, U, s, 3, j, k)) * std::conj(link_el(l, U, s_t, mu, k, i));
                                b[i][j] = v;  // (U_t(s) U_mu(s+t))^+ = U_mu(s+t)^+ U_t(s)^+
                            }
                        double tr = 0;
                        for (int i = 0; i < N; ++i)
                            for (int k = 0; k < N; ++k) tr += (a[i][k] * b[k][i]).real();
                        sum += tr;
                    }
                }
    return (double)(sum / (3.0L * N * (long double)l.V));
}

std::string num(int n) { return std::to_string(n); }

// ---------------------------------------------------------------- leonard_format
int write_leonard(const std::string& dir, const std::string& name, int number, const Lat& l, const double* U,
                  double plaq) {
    const int N = l.N;
    XdrOut o;
    o.buf.reserve((size_t)l.V * (16 + 4 * N * N * 16));
    for (int x = 0; x < l.L[0]; ++x)
        for (int y = 0; y < l.L[1]; ++y)
            for (int z = 0; z < l.L[2]; ++z)
                for (int t = 0; t < l.L[3]; ++t) {
                    const long s = l.lex(x, y, z, t);
                    o.i32(x);
                    o.i32(y);
                    o.i32(z);
                    o.i32(t);
                    for (int mu = 0; mu < 4; ++mu)
                        for (int i = 0; i < N; ++i)
                            for (int j = 0; j < N; ++j) {
                                const cd v = link_el(l, U, s, mu, i, j);
                                o.f64(v.real());
                                o.f64(v.imag());
                            }
                }
    const std::string base = dir + name + "_" + num(number);
    int rc = write_file(base + "_0.txt", o.buf.data(), o.buf.size());
    if (rc) return rc;
    char desc[512];
    // one process: processor grid 1 1 1 1, every site local; plaquette to 13 significant digits
    const int n = std::snprintf(desc, sizeof(desc), "%d %d %d %d\n1 1 1 1\n1\n%ld\n%.13g\n", l.L[0], l.L[1], l.L[2],
                                l.L[3], l.V, plaq);
    return write_file(base + ".descriptor.txt", desc, (size_t)n);
}

int read_leonard(const std::string& dir, const std::string& name, int number, const Lat& l, double* U,
                 double* plaq_read) {
    const int N = l.N;
    const std::string base = dir + name + "_" + num(number);
    std::vector<unsigned char> text;
    int rc = read_file(base + ".descriptor.txt", text);
    if (rc) return rc;
    text.push_back(0);
    long d[10];
    double plaq = 0;
    if (std::sscanf(reinterpret_cast<const char*>(text.data()), "%ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %lf", &d[0],
                    &d[1], &d[2], &d[3], &d[4], &d[5], &d[6], &d[7], &d[8], &d[9], &plaq) != 11)
        return fail(LYM_ERR_INVALID, "malformed descriptor %s.descriptor.txt", base.c_str());
    for (int m = 0; m < 4; ++m)
        if (d[m] != l.L[m])
            return fail(LYM_ERR_INVALID, "different lattice size in %s: %ld %ld %ld %ld, configured %d %d %d %d",
                        base.c_str(), d[0], d[1], d[2], d[3], l.L[0], l.L[1], l.L[2], l.L[3]);
    const long nproc = d[8], localsize = d[9];
    if (nproc < 1 || localsize < 0) return fail(LYM_ERR_INVALID, "malformed descriptor %s", base.c_str());
    std::vector<unsigned char> data;
    for (long p = 0; p < nproc; ++p) {
        const std::string path = base + "_" + num((int)p) + ".txt";
        if ((rc = read_file(path, data))) return rc;
        XdrIn in{data.data(), data.size(), 0};
        for (long r = 0; r < localsize; ++r) {
            int c[4];
            for (int m = 0; m < 4; ++m) c[m] = in.i32();
            bool inside = true;
            for (int m = 0; m < 4; ++m) inside = inside && c[m] >= 0 && c[m] < l.L[m];
            const long s = inside ? l.lex(c[0], c[1], c[2], c[3]) : -1;
            for (int mu = 0; mu < 4; ++mu)
                for (int i = 0; i < N; ++i)
                    for (int j = 0; j < N; ++j) {
                        const double re = in.f64(), im = in.f64();
                        if (s >= 0) set_el(l, U, s, mu, i, j, cd(re, im));
                    }
            if (!in.ok) return fail(LYM_ERR_INVALID, "%s ends after %ld of %ld records", path.c_str(), r, localsize);
        }
    }
    if (plaq_read) *plaq_read = plaq;
    return LYM_OK;
}

// ---------------------------------------------------------------- muenster_format
std::string muenster_path(const std::string& dir, const std::string& name, int number, int N) {
    return N == 2 ? dir + "0:" + num(number) + ":" + name : dir + name + ".config-" + num(number);
}

int write_muenster(const std::string& path, const Lat& l, const double* U, const lym_config_info& info,
                   double plaq) {
    const int N = l.N;
    XdrOut o;
    o.buf.reserve((size_t)l.V * 4 * (N == 2 ? 16 : 48) + 256);
    for (int t = 0; t < l.L[3]; ++t)
        for (int z = 0; z < l.L[2]; ++z)
            for (int y = 0; y < l.L[1]; ++y)
                for (int x = 0; x < l.L[0]; ++x) {
                    const long s = l.lex(x, y, z, t);
                    for (int mu = 0; mu < 4; ++mu) {
                        if (N == 2) {
                            // first row of U^+: conj(U00), conj(U10)
                            for (int j = 0; j < 2; ++j) {
                                const cd v = std::conj(link_el(l, U, s, mu, j, 0));
                                o.f32((float)v.real());
                                o.f32((float)v.imag());
                            }
                        } else {
                            // first two columns of U^*, rows outer
                            for (int i = 0; i < 3; ++i)
                                for (int j = 0; j < 2; ++j) {
                                    const cd v = link_el(l, U, s, mu, i, j);
                                    o.f32((float)v.real());
                                    o.f32((float)-v.imag());
                                }
                        }
                    }
                }
    if (N == 2) {
        // one kappa: no negated entries; alpha, epsilon, lambda and the
        // polynomial orders are written as zeros (OutputSweep.cpp:209-269)
        o.f32((float)plaq);
        o.f32(0.f);
        o.f32(0.f);
        o.f32((float)info.beta);
        o.f32((float)info.kappa);
        o.f32(0.f);            // alpha
        o.f32(0.f);            // epsilon
        o.f32(0.f);            // lambda
        o.i32(info.sweep);
        for (int k = 0; k < 5; ++k) o.i32(0);
        for (int m = 0; m < 4; ++m) o.i32(l.L[m]);
        o.i32(0);              // lvrc
    } else {
        for (int m = 0; m < 4; ++m) o.i32(l.L[m]);
        o.f64((double)(float)info.beta);  // the reference stores beta through a float
        o.f64(-1.0 / 12.0);
        o.f64(plaq);
    }
    return write_file(path, o.buf.data(), o.buf.size());
}

// reading: the SU(2) link from the first row (z0, z1) of U^+
// (ReadStartGaugeConfiguration.cpp:163-167): U00 = conj z0, U11 = z0, U01 = -z1, U10 = conj z1
void su2_from_row(const Lat& l, double* U, long s, int mu, cd z0, cd z1) {
    set_el(l, U, s, mu, 0, 0, std::conj(z0));
    set_el(l, U, s, mu, 1, 1, z0);
    set_el(l, U, s, mu, 0, 1, -z1);
    set_el(l, U, s, mu, 1, 0, std::conj(z1));
}

// SU(3) from two float-rounded columns: normalise col 0, orthogonalise and
// normalise col 1, col 2 = conj(col0 x col1) (ReadStartGaugeConfiguration.cpp:338-366)
void su3_from_columns(const Lat& l, double* U, long s, int mu, cd m[3][2]) {
    double n0 = 0;
    for (int i = 0; i < 3; ++i) n0 += std::norm(m[i][0]);
    const double r0 = 1.0 / std::sqrt(n0);
    for (int i = 0; i < 3; ++i) m[i][0] *= r0;
    cd d = 0;
    for (int i = 0; i < 3; ++i) d += m[i][1] * std::conj(m[i][0]);
    for (int i = 0; i < 3; ++i) m[i][1] -= d * m[i][0];
    double n1 = 0;
    for (int i = 0; i < 3; ++i) n1 += std::norm(m[i][1]);
    const double r1 = 1.0 / std::sqrt(n1);
    for (int i = 0; i < 3; ++i) m[i][1] *= r1;
    cd c2[3];
    c2[0] = std::conj(m[1][0]) * std::conj(m[2][1]) - std::conj(m[2][0]) * std::conj(m[1][1]);
    c2[1] = std::conj(m[2][0]) * std::conj(m[0][1]) - std::conj(m[0][0]) * std::conj(m[2][1]);
    c2[2] = std::conj(m[0][0]) * std::conj(m[1][1]) - std::conj(m[1][0]) * std::conj(m[0][1]);
    for (int i = 0; i < 3; ++i) {
        set_el(l, U, s, mu, i, 0, m[i][0]);
        set_el(l, U, s, mu, i, 1, m[i][1]);
        set_el(l, U, s, mu, i, 2, c2[i]);
    }
}

int read_muenster(const std::string& path, const Lat& l, double* U, lym_config_info* info) {
    const int N = l.N;
    std::vector<unsigned char> data;
    int rc = read_file(path, data);
    if (rc) return rc;
    XdrIn in{data.data(), data.size(), 0};
    for (int t = 0; t < l.L[3]; ++t)
        for (int z = 0; z < l.L[2]; ++z)
            for (int y = 0; y < l.L[1]; ++y)
                for (int x = 0; x < l.L[0]; ++x) {
                    const long s = l.lex(x, y, z, t);
                    for (int mu = 0; mu < 4; ++mu) {
                        if (N == 2) {
                            const float a = in.f32(), b = in.f32(), c = in.f32(), e = in.f32();
                            su2_from_row(l, U, s, mu, cd(a, b), cd(c, e));
                        } else {
                            cd m[3][2];
                            for (int i = 0; i < 3; ++i)
                                for (int j = 0; j < 2; ++j) {
                                    const float a = in.f32(), b = in.f32();
                                    m[i][j] = cd(a, -(double)b);
                                }
                            su3_from_columns(l, U, s, mu, m);
                        }
                    }
                }
    if (!in.ok) return fail(LYM_ERR_INVALID, "%s: read error for gauge fields (file too short)", path.c_str());
    lym_config_info got;
    std::memset(&got, 0, sizeof(got));
    int lat[4];
    if (N == 2) {
        got.plaquette = in.f32();
        (void)in.f32();  // correction factor
        (void)in.f32();  // smallest eigenvalue
        got.beta = in.f32();
        float kappa = -1.f;
        int nkappa = 0;
        while (kappa < 0.f && in.ok) {  // negative kappas precede the last, positive one
            kappa = in.f32();
            ++nkappa;
        }
        got.kappa = kappa;
        for (int k = 0; k < nkappa; ++k) (void)in.f32();      // alphas
        for (int k = 0; k < 2 * nkappa; ++k) (void)in.f32();  // epsilon, lambda pairs
        got.sweep = in.i32();
        for (int k = 0; k < 5 * nkappa; ++k) (void)in.i32();  // polynomial orders
        for (int m = 0; m < 4; ++m) lat[m] = in.i32();
        (void)in.i32();  // lvrc
        if (!in.ok) return fail(LYM_ERR_INVALID, "%s: truncated trailer", path.c_str());
    } else {
        for (int m = 0; m < 4; ++m) lat[m] = in.i32();
        got.beta = in.f64();
        got.c1 = in.f64();
        if (!in.ok) return fail(LYM_ERR_INVALID, "%s: truncated trailer", path.c_str());
        const double p = in.f64();
        // a missing plaquette is only a warning in the reference (:470-476)
        got.plaquette = in.ok ? p : NAN;
    }
    for (int m = 0; m < 4; ++m)
        if (lat[m] != l.L[m])
            return fail(LYM_ERR_INVALID, "%s: lattice size is different: %d %d %d %d, configured %d %d %d %d",
                        path.c_str(), lat[0], lat[1], lat[2], lat[3], l.L[0], l.L[1], l.L[2], l.L[3]);
    if (info) *info = got;
    return LYM_OK;
}

}  // namespace

extern "C" {

int lym_temporal_plaquette(const int dims[4], int ncolors, const double* links, double* out) {
    Lat l;
    int rc = make_lat(dims, ncolors, l);
    if (rc) return rc;
    if (!links || !out) return fail(LYM_ERR_INVALID, "null argument");
    *out = temporal_plaquette(l, links);
    return LYM_OK;
}

int lym_config_write(int format, const char* directory, const char* name, int number, const int dims[4],
                     int ncolors, const double* links, lym_config_info* info) {
    Lat l;
    int rc = make_lat(dims, ncolors, l);
    if (rc) return rc;
    if (!directory || !name || !links || !info) return fail(LYM_ERR_INVALID, "null argument");
    const double plaq = temporal_plaquette(l, links);
    if (format == LYM_FORMAT_LEONARD) {
        rc = write_leonard(directory, name, number, l, links, plaq);
    } else if (format == LYM_FORMAT_MUENSTER) {
        rc = write_muenster(muenster_path(directory, name, number, ncolors), l, links, *info, plaq);
        if (!rc && ncolors == 3) info->c1 = -1.0 / 12.0;
    } else {
        return fail(LYM_ERR_INVALID, "unknown configuration format %d", format);
    }
    if (!rc) info->plaquette = plaq;
    return rc;
}

int lym_config_read(int format, const char* directory, const char* name, int number, const int dims[4], int ncolors,
                    double* links, lym_config_info* info) {
    Lat l;
    int rc = make_lat(dims, ncolors, l);
    if (rc) return rc;
    if (!directory || !name || !links) return fail(LYM_ERR_INVALID, "null argument");
    if (format == LYM_FORMAT_LEONARD) {
        double plaq = 0;
        rc = read_leonard(directory, name, number, l, links, &plaq);
        if (!rc && info) {
            std::memset(info, 0, sizeof(*info));
            info->plaquette = plaq;
        }
        return rc;
    }
    if (format == LYM_FORMAT_MUENSTER) return read_muenster(muenster_path(directory, name, number, ncolors), l, links, info);
    return fail(LYM_ERR_INVALID, "unknown configuration format %d", format);
}

}  // extern "C"
