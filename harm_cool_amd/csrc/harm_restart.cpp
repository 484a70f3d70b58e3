// harm_restart.cpp -- the reference's ASCII restart file (restart.c:73-383).
//
// Layout, as restart_write() prints it: one header line of run parameters and
// counters ("%10d" integers, "%32.22e" doubles, restart.c:105-138), then one
// line per zone with its 8 primitives in (h, i, j) order over all zones
// including ghosts (restart.c:144-147), then on rank 0 the size of ener.out
// (restart.c:150-157), and with the tabulated EOS one temperature per zone in
// the same order (cooleos.c:996-1014).  "%32.22e" keeps 23 significant digits,
// so a double read back with strtod (fscanf "%lf", restart.c:329-374) is the
// double that was written.
//
// Every value is exactly 32 characters wide, so each zone line is 257 bytes;
// writing formats zone ranges on several threads straight into their place in
// the output, and reading parses fixed-width lines in parallel whenever the
// file has that shape (any other whitespace layout is read token by token, as
// fscanf would).
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "harm_host.hpp"

namespace harm_host {

namespace {

constexpr int W = 32;                  // "%32.22e"
constexpr size_t LINE = 8 * W + 1;     // one zone: 8 values and '\n'
constexpr size_t CHUNK = 1u << 20;     // zones formatted per pass

int nthreads()
{
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h ? h : 1u));
}

template <class F>
void parallel_for(size_t n, F f)
{
    const int nt = (int)std::min<size_t>((size_t)nthreads(), std::max<size_t>(n / 4096, 1));
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        const size_t a = n * t / nt, b = n * (t + 1) / nt;
        th.emplace_back([=] { f(a, b); });
    }
    for (auto &x : th) x.join();
}

void fmt_dbl(char *dst, double v)
{
    char tmp[64];
    const int len = snprintf(tmp, sizeof(tmp), "%32.22e", v);
    // every finite double, inf and nan fit the 32-character field
    memcpy(dst, tmp, (size_t)std::min(len, W));
}

std::string header_text(const harm_restart_hdr &h)
{
    std::string s;
    char b[64];
    auto I = [&](int32_t v) {
        snprintf(b, sizeof(b), "%10d", v);
        s += b;
    };
    auto D = [&](double v) {
        snprintf(b, sizeof(b), "%32.22e", v);
        s += b;
    };
    I(h.N1);
    I(h.N2);
    I(h.N3);
    D(h.t);
    D(h.tf);
    I(h.nstep);
    D(h.a);
    D(h.gam);
    D(h.cour);
    D(h.DTd);
    D(h.DTl);
    D(h.DTi);
    I(h.DTr);
    I(h.dump_cnt);
    I(h.image_cnt);
    I(h.rdump_cnt);
    D(h.dt);
    I(h.lim);
    I(h.failed);
    D(h.Rin);
    D(h.Rout);
    D(h.hslope);
    D(h.R0);
    D(h.mass_start);
    D(h.mass_in);
    D(h.mass_out);
    D(h.tdump);
    D(h.timage);
    D(h.tlog);
    D(h.defcon);
    I(h.nfailed);
    D(h.Rbr);
    D(h.cyl_R0);
    D(h.cyl_Th0);
    s += "\n";
    return s;
}

// Whitespace-separated tokens, as fscanf reads them.
struct Tokens {
    const char *p, *end;
    bool next(const char *&s, size_t &n)
    {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '\f' || *p == '\v')) p++;
        if (p >= end) return false;
        s = p;
        while (p < end && !(*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '\f' || *p == '\v')) p++;
        n = (size_t)(p - s);
        return true;
    }
    bool dbl(double &v)
    {
        const char *s;
        size_t n;
        if (!next(s, n) || n >= 64) return false;
        char b[64];
        memcpy(b, s, n);
        b[n] = 0;
        char *e;
        v = strtod(b, &e);
        return e == b + n;
    }
    bool i32(int32_t &v, bool *is_int = nullptr)
    {
        const char *s;
        size_t n;
        if (!next(s, n) || n >= 64) return false;
        char b[64];
        memcpy(b, s, n);
        b[n] = 0;
        char *e;
        errno = 0;
        const long x = strtol(b, &e, 10);
        const bool ok = e == b + n && errno == 0 && x >= INT32_MIN && x <= INT32_MAX;
        if (is_int) *is_int = ok;
        v = (int32_t)x;
        return ok;
    }
};

int read_header(Tokens &tk, harm_restart_hdr &h)
{
    memset(&h, 0, sizeof(h));
    bool ok = tk.i32(h.N1) && tk.i32(h.N2) && tk.i32(h.N3) && tk.dbl(h.t) && tk.dbl(h.tf) && tk.i32(h.nstep) &&
              tk.dbl(h.a) && tk.dbl(h.gam) && tk.dbl(h.cour) && tk.dbl(h.DTd) && tk.dbl(h.DTl) && tk.dbl(h.DTi) &&
              tk.i32(h.DTr) && tk.i32(h.dump_cnt) && tk.i32(h.image_cnt) && tk.i32(h.rdump_cnt) && tk.dbl(h.dt) &&
              tk.i32(h.lim) && tk.i32(h.failed) && tk.dbl(h.Rin) && tk.dbl(h.Rout) && tk.dbl(h.hslope) &&
              tk.dbl(h.R0) && tk.dbl(h.mass_start) && tk.dbl(h.mass_in) && tk.dbl(h.mass_out) && tk.dbl(h.tdump) &&
              tk.dbl(h.timage) && tk.dbl(h.tlog) && tk.dbl(h.defcon) && tk.i32(h.nfailed) && tk.dbl(h.Rbr) &&
              tk.dbl(h.cyl_R0) && tk.dbl(h.cyl_Th0);
    h.ener_size = -1;
    if (!ok) return set_error("restart file: malformed or truncated header", 3);
    if (h.N1 < 1 || h.N2 < 1 || h.N3 < 1 || (double)(h.N1 + 4) * (h.N2 + 4) * h.N3 >= 2147483648.0)
        return set_error("restart file: implausible grid size in the header", 3);
    return 0;
}

// The whole file, or its first max_bytes (the header is under 1 KB).
int slurp(const char *path, std::vector<char> &buf, size_t max_bytes = SIZE_MAX)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return set_error(std::string("cannot open restart file ") + path, 2);
    if (fseek(fp, 0, SEEK_END) != 0) {
        fclose(fp);
        return set_error("restart file: cannot seek", 2);
    }
    const long n = ftell(fp);
    rewind(fp);
    buf.resize(std::min(n > 0 ? (size_t)n : 0, max_bytes));
    const size_t got = buf.empty() ? 0 : fread(buf.data(), 1, buf.size(), fp);
    fclose(fp);
    if (got != buf.size()) return set_error("restart file: short read", 2);
    return 0;
}

}  // namespace

int restart_write_file(const char *path, const harm_restart_hdr &h, const double *prim, const double *temp,
                       const TracerBlock *tb)
{
    const int sn1 = h.N1 + 4, sn2 = h.N2 + 4, n3 = h.N3;
    const size_t nz = (size_t)sn1 * sn2 * n3;
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_error(std::string("cannot open restart file for writing: ") + path, 2);
    bool ok = true;
    const std::string hdr = header_text(h);
    ok = ok && fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size();
    std::vector<char> buf;
    // body: zone index q runs over (h, i, j) with j fastest (restart.c:144)
    for (size_t q0 = 0; ok && q0 < nz; q0 += CHUNK) {
        const size_t nq = std::min(CHUNK, nz - q0);
        buf.resize(nq * LINE);
        parallel_for(nq, [&](size_t a, size_t b) {
            for (size_t r = a; r < b; r++) {
                const size_t q = q0 + r;
                const size_t hh = q / ((size_t)sn1 * sn2), rem = q % ((size_t)sn1 * sn2);
                const size_t ii = rem / sn2, jj = rem % sn2;
                char *line = buf.data() + r * LINE;
                for (int k = 0; k < 8; k++)
                    fmt_dbl(line + k * W, prim[(((size_t)k * sn1 + ii) * sn2 + jj) * n3 + hh]);
                line[8 * W] = '\n';
            }
        });
        ok = fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    }
    if (ok && h.ener_size >= 0) ok = fprintf(fp, "%10d", h.ener_size) > 0;
    if (ok && tb) {
        // restart_write_tracers (tracers.c:565-606), no separators but the fields' own
        std::string t;
        char b[64];
        for (double v : tb->par) {
            snprintf(b, sizeof(b), "%32.22e", v);
            t += b;
        }
        snprintf(b, sizeof(b), "%10d", (int)tb->i.size());
        t += b;
        for (size_t q = 0; q < tb->i.size(); q++) {
            snprintf(b, sizeof(b), "%10d%10d%10d", tb->i[q], tb->j[q], tb->h[q]);
            t += b;
            for (int k = 0; k < 4; k++) {
                snprintf(b, sizeof(b), "%32.22e", tb->X[4 * q + k]);
                t += b;
            }
            snprintf(b, sizeof(b), "%10d", tb->size[q]);
            t += b;
        }
        ok = fwrite(t.data(), 1, t.size(), fp) == t.size();
    }
    if (ok && temp) {
        for (size_t q0 = 0; ok && q0 < nz; q0 += CHUNK) {
            const size_t nq = std::min(CHUNK, nz - q0);
            buf.resize(nq * W);
            parallel_for(nq, [&](size_t a, size_t b) {
                for (size_t r = a; r < b; r++) {
                    const size_t q = q0 + r;
                    const size_t hh = q / ((size_t)sn1 * sn2), rem = q % ((size_t)sn1 * sn2);
                    const size_t ii = rem / sn2, jj = rem % sn2;
                    fmt_dbl(buf.data() + r * W, temp[(ii * sn2 + jj) * n3 + hh]);
                }
            });
            ok = fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
        }
    }
    if (fclose(fp) != 0) ok = false;
    return ok ? 0 : set_error(std::string("error writing restart file ") + path, 2);
}

std::string dump_header_text(const harm_restart_hdr &h, const double *startx, const double *dx)
{
    // dump.c:1745-1790: 2D files name N2 and the theta origin and spacing, 3D
    // files N3 and the phi ones
    std::string s;
    char b[64];
    auto I = [&](int32_t v) {
        snprintf(b, sizeof(b), "%10d", v);
        s += b;
    };
    auto D = [&](double v) {
        snprintf(b, sizeof(b), "%32.22e", v);
        s += b;
    };
    const bool d3 = h.N3 > 1;
    D(h.t);
    I(h.N1);
    I(d3 ? h.N3 : h.N2);
    D(startx[1]);
    D(d3 ? startx[3] : startx[2]);
    D(dx[1]);
    D(d3 ? dx[3] : dx[2]);
    D(h.tf);
    I(h.nstep);
    D(h.a);
    D(h.gam);
    D(h.cour);
    D(h.DTd);
    D(h.DTl);
    D(h.DTi);
    I(h.DTr);
    I(h.dump_cnt);
    I(h.image_cnt);
    I(h.rdump_cnt);
    D(h.dt);
    I(h.lim);
    I(h.failed);
    D(h.Rin);
    D(h.Rout);
    D(h.hslope);
    D(h.R0);
    s += "\n";
    return s;
}

int dump_write_file(const char *path, const std::string &header, size_t nzones, int ncol, int stride,
                    const std::function<int(size_t, size_t, double *)> &fill)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_error(std::string("cannot open dump file for writing: ") + path, 2);
    bool ok = fwrite(header.data(), 1, header.size(), fp) == header.size();
    const size_t line = (size_t)ncol * W + 1, chunk = CHUNK / 4;
    std::vector<double> cols;
    std::vector<char> buf;
    for (size_t z0 = 0; ok && z0 < nzones; z0 += chunk) {
        const size_t nq = std::min(chunk, nzones - z0);
        cols.resize(nq * stride);
        int rc = fill(z0, nq, cols.data());
        if (rc) {
            fclose(fp);
            return rc;
        }
        buf.resize(nq * line);
        parallel_for(nq, [&](size_t a, size_t b) {
            for (size_t r = a; r < b; r++) {
                char *ln = buf.data() + r * line;
                for (int k = 0; k < ncol; k++) fmt_dbl(ln + k * W, cols[r * stride + k]);
                ln[ncol * W] = '\n';
            }
        });
        ok = fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    }
    if (fclose(fp) != 0) ok = false;
    return ok ? 0 : set_error(std::string("error writing dump file ") + path, 2);
}

int restart_read_file(const char *path, harm_restart_hdr &h, double *prim, double *temp, int *has_temp,
                      TracerBlock *tb, int *has_tb)
{
    if (has_tb) *has_tb = 0;
    std::vector<char> buf;
    int rc = slurp(path, buf, prim ? SIZE_MAX : (size_t)65536);
    if (rc) return rc;
    Tokens tk{buf.data(), buf.data() + buf.size()};
    if ((rc = read_header(tk, h))) return rc;
    if (has_temp) *has_temp = 0;
    if (!prim) return 0;
    const int sn1 = h.N1 + 4, sn2 = h.N2 + 4, n3 = h.N3;
    const size_t nz = (size_t)sn1 * sn2 * n3;
    auto put = [&](size_t q, int k, double v) {
        const size_t hh = q / ((size_t)sn1 * sn2), rem = q % ((size_t)sn1 * sn2);
        const size_t ii = rem / sn2, jj = rem % sn2;
        prim[(((size_t)k * sn1 + ii) * sn2 + jj) * n3 + hh] = v;
    };
    // fixed-width body: the header's '\n', then nz lines of exactly LINE bytes
    const char *body = tk.p;
    while (body < tk.end && *body != '\n') body++;
    bool fixed = body < tk.end && (size_t)(tk.end - (body + 1)) >= nz * LINE;
    if (fixed) {
        body++;
        for (size_t q = 0; q < nz && fixed; q += std::max<size_t>(nz / 64, 1)) fixed = body[q * LINE + 8 * W] == '\n';
        fixed = fixed && body[(nz - 1) * LINE + 8 * W] == '\n';
    }
    bool ok = true;
    if (fixed) {
        std::atomic<int> bad{0};
        parallel_for(nz, [&](size_t a, size_t b) {
            char f[W + 1];
            f[W] = 0;
            bool good = true;
            for (size_t q = a; q < b && good; q++) {
                const char *line = body + q * LINE;
                good = line[8 * W] == '\n';
                for (int k = 0; k < 8 && good; k++) {
                    memcpy(f, line + k * W, W);
                    char *e;
                    const double v = strtod(f, &e);
                    good = e != f && (*e == 0);
                    put(q, k, v);
                }
            }
            if (!good) bad.store(1);
        });
        ok = bad.load() == 0;
        if (ok) tk.p = body + nz * LINE;
    }
    if (!fixed || !ok) {
        // token by token from the start of the body, as fscanf("%lf") would
        Tokens t2{buf.data(), buf.data() + buf.size()};
        harm_restart_hdr h2;
        if ((rc = read_header(t2, h2))) return rc;
        for (size_t q = 0; q < nz; q++)
            for (int k = 0; k < 8; k++) {
                double v;
                if (!t2.dbl(v)) return set_error("restart file: body truncated or malformed", 3);
                put(q, k, v);
            }
        tk = t2;
    }
    // optional ener.out size (an integer token), then optional temperatures
    Tokens save = tk;
    int32_t es;
    bool is_int = false;
    if (tk.i32(es, &is_int) && is_int)
        h.ener_size = es;
    else
        tk = save;
    // a tracer block: eight doubles then an integer count (a temperature block
    // holds "%32.22e" values only, whose tokens always carry an exponent)
    {
        Tokens probe = tk;
        double par[8];
        bool blk = true;
        for (int k = 0; k < 8 && blk; k++) blk = probe.dbl(par[k]);
        int32_t cnt = 0;
        bool is_int = false;
        blk = blk && probe.i32(cnt, &is_int) && is_int && cnt >= 0;
        if (blk) {
            TracerBlock local;
            TracerBlock &T = tb ? *tb : local;
            std::copy(par, par + 8, T.par);
            T.i.resize(cnt);
            T.j.resize(cnt);
            T.h.resize(cnt);
            T.size.resize(cnt);
            T.X.resize(4 * (size_t)cnt);
            for (int32_t q = 0; q < cnt; q++) {
                bool good = probe.i32(T.i[q]) && probe.i32(T.j[q]) && probe.i32(T.h[q]);
                for (int k = 0; k < 4 && good; k++) good = probe.dbl(T.X[4 * (size_t)q + k]);
                good = good && probe.i32(T.size[q]);
                if (!good) return set_error("restart file: tracer block truncated or malformed", 3);
            }
            tk = probe;
            if (has_tb) *has_tb = 1;
        }
    }
    save = tk;
    const char *s;
    size_t n;
    if (temp && tk.next(s, n)) {
        tk = save;
        for (size_t q = 0; q < nz; q++) {
            double v;
            if (!tk.dbl(v)) return set_error("restart file: temperature block truncated or malformed", 3);
            const size_t hh = q / ((size_t)sn1 * sn2), rem = q % ((size_t)sn1 * sn2);
            temp[rem * n3 + hh] = v;
        }
        if (has_temp) *has_temp = 1;
    }
    return 0;
}

}  // namespace harm_host

extern "C" {

int harm_restart_write_host(const char *path, const harm_restart_hdr *h, const double *prim, const double *temp)
{
    if (!path || !h || !prim) return harm_host::set_error("null argument", 1);
    if (h->N1 < 1 || h->N2 < 1 || h->N3 < 1) return harm_host::set_error("restart header: bad grid size", 1);
    return harm_host::restart_write_file(path, *h, prim, temp);
}

int harm_restart_read_host(const char *path, harm_restart_hdr *h, double *prim, double *temp, int *has_temp)
{
    if (!path || !h) return harm_host::set_error("null argument", 1);
    return harm_host::restart_read_file(path, *h, prim, temp, has_temp);
}

int harm_restart_peek(const char *path, harm_restart_hdr *h)
{
    if (!path || !h) return harm_host::set_error("null argument", 1);
    return harm_host::restart_read_file(path, *h, nullptr, nullptr, nullptr);
}

}  // extern "C"
