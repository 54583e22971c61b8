// snapshot.cpp -- PeTar snapshot files (pth_snapshot_* of include/petar_host.h).
//
// A snapshot is what PeTar's ParticleSystem writes with writeParticleAscii /
// writeParticleBinary (src/petar.hpp:1905-1908, read back at
// src/petar.hpp:2757-2763): a FileHeader (src/io.hpp:210-282) followed by one
// FPSoft record per particle.  In the default build a record is
//   ParticleBase  mass, pos[3], vel[3], binary_state    (src/particle_base.hpp:23-26,178-193)
//   Ptcl          r_search, id, group data (2 x int64), changeover r_in, r_out
//                                                      (src/ptcl.hpp:38-41,158-163; src/changeover.hpp:141-150)
//   FPSoft        acc[3], pot_tot, pot_soft, n_ngb      (src/soft_ptcl.hpp:26-40,114-121)
// i.e. 20 eight-byte fields; the column names are those of PeTar's own
// reader (tools/analysis/data.py:390-399: mass_bk and status are the two
// group-data words).  Build options add fields:
//   PTH_SNAP_EXTERNAL  an external-potential build (RECORD_CM_IN_HEADER and
//                      EXTERNAL_POT_IN_PTCL, Makefile.in:90-93): the header
//                      carries pos_offset[3], vel_offset[3]; pot_ext follows
//                      pot_soft;
//   PTH_SNAP_SE        STELLAR_EVOLUTION (without BSE): radius, dm,
//                      time_record, time_interrupt follow binary_state.
// Binary files are the raw little-endian records (the header's 3 or 9
// words, then the records); ASCII files hold the header on the first line
// and one record per line, as the fprintf formats of the writeAscii members
// (%26.17e for reals, %lld for integers).
#include <omp.h>

#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/petar_host.h"

namespace pth {
int host_fail(int code, const std::string &msg);
}
using pth::host_fail;

namespace {

struct Field {
    const char *name;
    int width;    // 1 or 3
    bool integer; // int64 (else FP64)
};

// record layout for the given flags, in file order
std::vector<Field> layout(int flags)
{
    std::vector<Field> f = {{"mass", 1, false}, {"pos", 3, false}, {"vel", 3, false}, {"binary_state", 1, true}};
    if (flags & PTH_SNAP_SE) {
        f.push_back({"radius", 1, false});
        f.push_back({"dm", 1, false});
        f.push_back({"time_record", 1, false});
        f.push_back({"time_interrupt", 1, false});
    }
    f.push_back({"r_search", 1, false});
    f.push_back({"id", 1, true});
    f.push_back({"mass_bk", 1, true});
    f.push_back({"status", 1, true});
    f.push_back({"r_in", 1, false});
    f.push_back({"r_out", 1, false});
    f.push_back({"acc_soft", 3, false});
    f.push_back({"pot", 1, false});
    f.push_back({"pot_soft", 1, false});
    if (flags & PTH_SNAP_EXTERNAL) f.push_back({"pot_ext", 1, false});
    f.push_back({"n_nb", 1, true});
    return f;
}

int words(const std::vector<Field> &f)
{
    int w = 0;
    for (const Field &x : f) w += x.width;
    return w;
}

// parses one whitespace-separated number at p (skipping leading blanks);
// returns the position after it or nullptr
const char *parse_real(const char *p, const char *end, double &v)
{
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    if (p < end && *p == '+') p++;
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return nullptr;
    return r.ptr;
}

const char *parse_int(const char *p, const char *end, long long &v)
{
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    if (p < end && *p == '+') p++;
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return nullptr;
    return r.ptr;
}

}  // namespace

struct pth_snapshot {
    int flags = 0;
    long long file_id = 0, n = 0;
    double time = 0, pos_offset[3] = {0, 0, 0}, vel_offset[3] = {0, 0, 0};
    std::vector<Field> fields;
    int w = 0;                   // 8-byte words per record
    std::vector<uint64_t> rec;   // n records of w words (int64 or FP64 bit patterns)
};

extern "C" {

int pth_snapshot_new(long long n, int flags, pth_snapshot **out)
{
    if (!out) return host_fail(-1, "pth_snapshot_new: null output");
    *out = nullptr;
    if (n < 0) return host_fail(-1, "pth_snapshot_new: negative n");
    if (flags & ~(PTH_SNAP_EXTERNAL | PTH_SNAP_SE)) return host_fail(-1, "pth_snapshot_new: unknown flags");
    pth_snapshot *s = new pth_snapshot;
    s->flags = flags;
    s->n = n;
    s->fields = layout(flags);
    s->w = words(s->fields);
    s->rec.assign((size_t)n * s->w, 0);
    *out = s;
    return 0;
}

int pth_snapshot_read(const char *path, int format, int flags, pth_snapshot **out)
{
    if (!out) return host_fail(-1, "pth_snapshot_read: null output");
    *out = nullptr;
    if (!path) return host_fail(-1, "pth_snapshot_read: null path");
    if (format != PTH_SNAP_BINARY && format != PTH_SNAP_ASCII)
        return host_fail(-1, "pth_snapshot_read: format must be PTH_SNAP_BINARY or PTH_SNAP_ASCII");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return host_fail(-1, std::string("pth_snapshot_read: cannot open ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::vector<char> buf((size_t)std::max(size, 0LL));
    const size_t got = size > 0 ? std::fread(buf.data(), 1, (size_t)size, fp) : 0;
    std::fclose(fp);
    if ((long long)got != size) return host_fail(-1, std::string("pth_snapshot_read: short read of ") + path);
    pth_snapshot *s = nullptr;
    if (pth_snapshot_new(0, flags, &s)) return -1;
    const bool ext = (flags & PTH_SNAP_EXTERNAL) != 0;
    const int hw = ext ? 9 : 3;
    auto fail = [&](const std::string &m) {
        delete s;
        return host_fail(-1, "pth_snapshot_read: " + std::string(path) + ": " + m);
    };
    if (format == PTH_SNAP_BINARY) {
        if (size < 8 * hw) return fail("shorter than the header");
        std::memcpy(&s->file_id, buf.data(), 8);
        std::memcpy(&s->n, buf.data() + 8, 8);
        std::memcpy(&s->time, buf.data() + 16, 8);
        if (ext) {
            std::memcpy(s->pos_offset, buf.data() + 24, 24);
            std::memcpy(s->vel_offset, buf.data() + 48, 24);
        }
        if (s->n < 0) return fail("negative particle count in the header");
        const long long want = 8LL * hw + 8LL * s->w * s->n;
        if (size != want)
            return fail("size " + std::to_string(size) + " bytes, the header's " + std::to_string(s->n) +
                        " records of " + std::to_string(8 * s->w) + " bytes need " + std::to_string(want) +
                        " (check the build flags)");
        s->rec.resize((size_t)s->n * s->w);
        if (s->n) std::memcpy(s->rec.data(), buf.data() + 8 * hw, (size_t)(8LL * s->w * s->n));
        *out = s;
        return 0;
    }
    // ASCII: header line, then one record per line
    const char *p = buf.data(), *end = buf.data() + buf.size();
    const char *eol = (const char *)std::memchr(p, '\n', end - p);
    if (!eol) eol = end;
    {
        const char *q = p;
        if (!(q = parse_int(q, eol, s->file_id)) || !(q = parse_int(q, eol, s->n)) ||
            !(q = parse_real(q, eol, s->time)))
            return fail("cannot read the header (file id, N, time)");
        if (ext)
            for (int d = 0; d < 6 && q; d++) q = parse_real(q, eol, d < 3 ? s->pos_offset[d] : s->vel_offset[d - 3]);
        if (!q) return fail("cannot read the header offsets (external-mode layout)");
        while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
        if (q != eol) return fail("extra items on the header line (check the build flags)");
    }
    if (s->n < 0) return fail("negative particle count in the header");
    // line starts
    std::vector<const char *> line;
    line.reserve((size_t)s->n + 1);
    for (const char *q = eol < end ? eol + 1 : end; q < end;) {
        const char *e = (const char *)std::memchr(q, '\n', end - q);
        if (!e) e = end;
        const char *t = q;
        while (t < e && (*t == ' ' || *t == '\t' || *t == '\r')) t++;
        if (t < e) line.push_back(q);  // blank lines are skipped
        q = e + 1;
    }
    if ((long long)line.size() != s->n)
        return fail(std::to_string(line.size()) + " records for a header count of " + std::to_string(s->n));
    s->rec.resize((size_t)s->n * s->w);
    long long bad = -1;
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < s->n; k++) {
        const char *q = line[k];
        const char *e = (const char *)std::memchr(q, '\n', end - q);
        if (!e) e = end;
        uint64_t *r = s->rec.data() + (size_t)k * s->w;
        int o = 0;
        for (const Field &f : s->fields)
            for (int c = 0; c < f.width && q; c++, o++) {
                if (f.integer) {
                    long long v = 0;
                    q = parse_int(q, e, v);
                    std::memcpy(&r[o], &v, 8);
                } else {
                    double v = 0;
                    q = parse_real(q, e, v);
                    std::memcpy(&r[o], &v, 8);
                }
            }
        if (q) {
            while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
            if (q != e) q = nullptr;
        }
        if (!q) {
#pragma omp critical
            if (bad < 0 || k < bad) bad = k;
        }
    }
    if (bad >= 0)
        return fail("record " + std::to_string(bad) + " does not hold " + std::to_string(s->w) +
                    " numbers of the layout (check the build flags)");
    *out = s;
    return 0;
}

int pth_snapshot_header(const pth_snapshot *s, long long *file_id, long long *n, double *time, double *pos_offset,
                        double *vel_offset)
{
    if (!s) return host_fail(-1, "pth_snapshot_header: null snapshot");
    if (file_id) *file_id = s->file_id;
    if (n) *n = s->n;
    if (time) *time = s->time;
    if (pos_offset) std::memcpy(pos_offset, s->pos_offset, 24);
    if (vel_offset) std::memcpy(vel_offset, s->vel_offset, 24);
    return 0;
}

int pth_snapshot_set_header(pth_snapshot *s, long long file_id, double time, const double *pos_offset,
                            const double *vel_offset)
{
    if (!s) return host_fail(-1, "pth_snapshot_set_header: null snapshot");
    s->file_id = file_id;
    s->time = time;
    if (pos_offset) std::memcpy(s->pos_offset, pos_offset, 24);
    if (vel_offset) std::memcpy(s->vel_offset, vel_offset, 24);
    return 0;
}

// width of the named column (1 or 3), 0 if the layout has no such column
int pth_snapshot_column_width(const pth_snapshot *s, const char *name)
{
    if (!s || !name) return 0;
    for (const Field &f : s->fields)
        if (std::strcmp(f.name, name) == 0) return f.width;
    return 0;
}

static int column_copy(pth_snapshot *s, const char *name, void *io, bool to_user, const char *who)
{
    if (!s || !name || !io) return host_fail(-1, std::string(who) + ": null argument");
    int o = 0;
    for (const Field &f : s->fields) {
        if (std::strcmp(f.name, name) == 0) {
            uint64_t *u = static_cast<uint64_t *>(io);
            for (long long k = 0; k < s->n; k++)
                for (int c = 0; c < f.width; c++) {
                    uint64_t &r = s->rec[(size_t)k * s->w + o + c];
                    if (to_user) u[k * f.width + c] = r;
                    else r = u[k * f.width + c];
                }
            return 0;
        }
        o += f.width;
    }
    return host_fail(-1, std::string(who) + ": no column '" + name + "' in this layout");
}

int pth_snapshot_column(const pth_snapshot *s, const char *name, void *out)
{
    return column_copy(const_cast<pth_snapshot *>(s), name, out, true, "pth_snapshot_column");
}

int pth_snapshot_set_column(pth_snapshot *s, const char *name, const void *in)
{
    return column_copy(s, name, const_cast<void *>(in), false, "pth_snapshot_set_column");
}

int pth_snapshot_write(const pth_snapshot *s, const char *path, int format)
{
    if (!s || !path) return host_fail(-1, "pth_snapshot_write: null argument");
    if (format != PTH_SNAP_BINARY && format != PTH_SNAP_ASCII)
        return host_fail(-1, "pth_snapshot_write: format must be PTH_SNAP_BINARY or PTH_SNAP_ASCII");
    FILE *fp = std::fopen(path, format == PTH_SNAP_BINARY ? "wb" : "w");
    if (!fp) return host_fail(-1, std::string("pth_snapshot_write: cannot create ") + path);
    const bool ext = (s->flags & PTH_SNAP_EXTERNAL) != 0;
    bool ok = true;
    if (format == PTH_SNAP_BINARY) {
        ok = std::fwrite(&s->file_id, 8, 1, fp) == 1 && std::fwrite(&s->n, 8, 1, fp) == 1 &&
             std::fwrite(&s->time, 8, 1, fp) == 1;
        if (ext) ok = ok && std::fwrite(s->pos_offset, 8, 3, fp) == 3 && std::fwrite(s->vel_offset, 8, 3, fp) == 3;
        if (s->n) ok = ok && std::fwrite(s->rec.data(), 8 * s->w, (size_t)s->n, fp) == (size_t)s->n;
    } else {
        if (ext)
            ok = std::fprintf(fp, "%lld %lld %26.17e %26.17e %26.17e %26.17e %26.17e %26.17e %26.17e\n", s->file_id,
                              s->n, s->time, s->pos_offset[0], s->pos_offset[1], s->pos_offset[2], s->vel_offset[0],
                              s->vel_offset[1], s->vel_offset[2]) > 0;
        else
            ok = std::fprintf(fp, "%lld %lld %26.17e\n", s->file_id, s->n, s->time) > 0;
        for (long long k = 0; k < s->n && ok; k++) {
            const uint64_t *r = s->rec.data() + (size_t)k * s->w;
            int o = 0;
            for (const Field &f : s->fields)
                for (int c = 0; c < f.width; c++, o++) {
                    if (f.integer) {
                        long long v;
                        std::memcpy(&v, &r[o], 8);
                        std::fprintf(fp, "%lld ", v);
                    } else {
                        double v;
                        std::memcpy(&v, &r[o], 8);
                        std::fprintf(fp, "%26.17e ", v);
                    }
                }
            ok = std::fputc('\n', fp) != EOF;
        }
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return host_fail(-1, std::string("pth_snapshot_write: write error on ") + path);
    return 0;
}

void pth_snapshot_free(pth_snapshot *s) { delete s; }

}  // extern "C"
