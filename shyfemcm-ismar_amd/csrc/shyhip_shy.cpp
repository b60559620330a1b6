// shyhip_shy.cpp -- the SHY output format (src/utils/shyutil/io_shy.f90),
// written from host arrays or from a model's device state.
//
// A SHY file is a Fortran unformatted sequential file: every record is its
// byte count (4-byte little-endian), the payload, and the count again.  The
// header (shy_write_header, io_shy.f90:1550-1606, format version 14) is
//   idshy nvers | ftype | nkn nel npr nlv nvar | nfix | date time | nsimpar |
//   simpar(nsimpar) | title (80 chars) | femver (80 chars) | (empty) |
//   nen3v(3,nel) | ipev | ipv | iarv | iarnv | xgv | ygv | hm3v(3,nel) |
//   hlv | ilhv | ilhkv | (empty)
// and each data record (shy_write_record, :1610-1681) is two records,
//   dtime ivar n m lmax
//   values
// with the values of a 3D record packed node by node (element by element)
// over the node's (element's) own layers (vals2linear, to_linear.f90:76).
// A hydro time record (shy_write_hydro_records, shy_util.f90:1210-1245) is
// four data records: znv (ivar 1, nodes), zenv (ivar 1, elements, m = 3),
// utlnv and vtlnv (ivar 3, elements, nlv layers).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shyhip_internal.h"

struct shyhip_shyfile {
    FILE *fp = nullptr;
    int ftype = 0, nkn = 0, nel = 0, nlv = 0, nvar = 0;
    std::vector<int> ilhv, ilhkv;
    std::string err;
};

namespace {

constexpr int kIdShy = 1617, kVersion = 14, kNSimpar = 4;

// one record from a list of (pointer, bytes) pieces
struct Piece {
    const void *p;
    size_t n;
};

int put_record(FILE *fp, std::string &err, std::initializer_list<Piece> parts) {
    size_t total = 0;
    for (const Piece &q : parts) total += q.n;
    if (total > 0x7fffffffu) {
        err = "record larger than 2 GiB";
        return 1;
    }
    const int32_t len = (int32_t)total;
    bool ok = fwrite(&len, 4, 1, fp) == 1;
    for (const Piece &q : parts)
        if (q.n) ok = ok && fwrite(q.p, 1, q.n, fp) == q.n;
    ok = ok && fwrite(&len, 4, 1, fp) == 1;
    if (!ok) err = "write error";
    return ok ? 0 : 1;
}

int put_record(shyhip_shyfile *f, std::initializer_list<Piece> parts) { return put_record(f->fp, f->err, parts); }

void text80(char *dst, const char *src) {
    std::memset(dst, ' ', 80);
    if (src) std::memcpy(dst, src, std::min<size_t>(80, std::strlen(src)));
}

}  // namespace

extern "C" {

int shyhip_shy_open(const char *path, int ftype, int nkn, int nel, int npr, int nlv, int nvar, int date, int time,
                    const char *title, const char *femver, const float *simpar, const int *nen3v, const int *ipev,
                    const int *ipv, const int *iarv, const int *iarnv, const float *xgv, const float *ygv,
                    const float *hm3v, const float *hlv, const int *ilhv, const int *ilhkv, shyhip_shyfile **out) {
    if (!out) return 1;
    *out = nullptr;
    if (!path || nkn <= 0 || nel <= 0 || nlv <= 0 || nvar <= 0 || !nen3v || !xgv || !ygv || !hm3v || !hlv || !ilhv ||
        !ilhkv)
        return 1;
    auto *f = new shyhip_shyfile;
    f->fp = std::fopen(path, "wb");
    if (!f->fp) {
        delete f;
        return 1;
    }
    f->ftype = ftype; f->nkn = nkn; f->nel = nel; f->nlv = nlv; f->nvar = nvar;
    f->ilhv.assign(ilhv, ilhv + nel);
    f->ilhkv.assign(ilhkv, ilhkv + nkn);
    std::vector<int> seq_n(nkn), seq_e(nel), zero_n(nkn, 0), zero_e(nel, 0);
    for (int i = 0; i < nkn; i++) seq_n[i] = i + 1;
    for (int i = 0; i < nel; i++) seq_e[i] = i + 1;
    const int hdr0[2] = {kIdShy, kVersion}, dims[5] = {nkn, nel, npr, nlv, nvar}, nfix = 0;
    const int dt[2] = {date, time}, nsim = kNSimpar;
    float sp[kNSimpar] = {0.f, 0.f, 0.f, 0.f};
    if (simpar) std::memcpy(sp, simpar, sizeof sp);
    char t80[80], v80[80];
    text80(t80, title);
    text80(v80, femver);
    const size_t I = sizeof(int), F = sizeof(float);
    int rc = put_record(f, {{hdr0, 2 * I}}) || put_record(f, {{&ftype, I}}) || put_record(f, {{dims, 5 * I}}) ||
             put_record(f, {{&nfix, I}}) || put_record(f, {{dt, 2 * I}}) || put_record(f, {{&nsim, I}}) ||
             put_record(f, {{sp, kNSimpar * F}}) || put_record(f, {{t80, 80}}) || put_record(f, {{v80, 80}}) ||
             put_record(f, {}) || put_record(f, {{nen3v, 3 * (size_t)nel * I}}) ||
             put_record(f, {{ipev ? ipev : seq_e.data(), nel * I}}) ||
             put_record(f, {{ipv ? ipv : seq_n.data(), nkn * I}}) ||
             put_record(f, {{iarv ? iarv : zero_e.data(), nel * I}}) ||
             put_record(f, {{iarnv ? iarnv : zero_n.data(), nkn * I}}) || put_record(f, {{xgv, nkn * F}}) ||
             put_record(f, {{ygv, nkn * F}}) || put_record(f, {{hm3v, 3 * (size_t)nel * F}}) ||
             put_record(f, {{hlv, nlv * F}}) || put_record(f, {{ilhv, nel * I}}) ||
             put_record(f, {{ilhkv, nkn * I}}) || put_record(f, {});
    if (rc) {
        std::fclose(f->fp);
        delete f;
        return 1;
    }
    *out = f;
    return 0;
}

// one data record: c(nlvddi, n*m) in the reference's layout
int shyhip_shy_write_record(shyhip_shyfile *f, double dtime, int ivar, int belem, int n, int m, int lmax, int nlvddi,
                            const float *c) {
    if (!f || !f->fp || !c || n <= 0 || m <= 0 || lmax <= 0 || nlvddi < lmax) return 1;
    if (n != (belem ? f->nel : f->nkn)) {
        f->err = "shy: n must be nel (elements) or nkn (nodes)";
        return 1;
    }
    struct {
        double dtime;
        int32_t ivar, n, m, lmax;
    } __attribute__((packed)) head{dtime, ivar, n, m, lmax};
    std::vector<float> v;
    if (lmax == 1) {
        v.resize((size_t)n * m);
        for (size_t i = 0; i < v.size(); i++) v[i] = c[i * nlvddi];
    } else {
        if (m != 1 || lmax != nlvddi) {
            f->err = "shy: a 3D record needs m = 1 and lmax = nlvddi";
            return 1;
        }
        const std::vector<int> &il = belem ? f->ilhv : f->ilhkv;
        for (int j = 0; j < n; j++) {
            const int l1 = std::min(nlvddi, il[j]);
            for (int l = 0; l < l1; l++) v.push_back(c[(size_t)j * nlvddi + l]);
        }
    }
    // the record line and the values are two Fortran records (two writes)
    return put_record(f, {{&head, sizeof head}}) || put_record(f, {{v.data(), v.size() * sizeof(float)}});
}

int shyhip_shy_close(shyhip_shyfile *f) {
    if (!f) return 0;
    const int rc = (f->fp && std::fclose(f->fp) != 0) ? 1 : 0;
    delete f;
    return rc;
}

const char *shyhip_shy_error(const shyhip_shyfile *f) { return f ? f->err.c_str() : "null file"; }

// shy_write_hydro_records from the device state (znv, zenv, utlnv, vtlnv)
int shyhip_shy_write_hydro(shyhip_model *h, shyhip_shyfile *f, double dtime) {
    if (!h || !f) return 1;
    shy::Model &M = h->M;
    if (f->nkn != M.nkn || f->nel != M.nel || f->nlv != M.nlvdi) {
        f->err = "shy: file and model sizes differ";
        return 1;
    }
    const int L = M.nlvdi;
    std::vector<float> z(M.nkn), ze(3 * (size_t)M.nel), u((size_t)L * M.nel), v((size_t)L * M.nel);
    if (shyhip_pull(h, SHYHIP_ZNV, z.data(), SHYHIP_F32, z.size()) ||
        shyhip_pull(h, SHYHIP_ZENV, ze.data(), SHYHIP_F32, ze.size()) ||
        shyhip_pull(h, SHYHIP_UTLNV, u.data(), SHYHIP_F32, u.size()) ||
        shyhip_pull(h, SHYHIP_VTLNV, v.data(), SHYHIP_F32, v.size()))
        return 1;
    return shyhip_shy_write_record(f, dtime, 1, 0, M.nkn, 1, 1, 1, z.data()) ||
           shyhip_shy_write_record(f, dtime, 1, 1, M.nel, 3, 1, 1, ze.data()) ||
           shyhip_shy_write_record(f, dtime, 3, 1, M.nel, 1, L, L, u.data()) ||
           shyhip_shy_write_record(f, dtime, 3, 1, M.nel, 1, L, L, v.data());
}

// one node 3D field of the device state (a tracer, SHYHIP_TRACER0 + itr)
// as a scalar record (shy_write_scalar_record)
int shyhip_shy_write_node3d(shyhip_model *h, shyhip_shyfile *f, double dtime, int ivar, int field) {
    if (!h || !f) return 1;
    shy::Model &M = h->M;
    if (f->nkn != M.nkn || f->nlv != M.nlvdi) {
        f->err = "shy: file and model sizes differ";
        return 1;
    }
    std::vector<float> c((size_t)M.nlvdi * M.nkn);
    if (shyhip_pull(h, field, c.data(), SHYHIP_F32, c.size())) return 1;
    return shyhip_shy_write_record(f, dtime, ivar, 0, M.nkn, 1, M.nlvdi, M.nlvdi, c.data());
}

}  // extern "C"

// ---------------------------------------------------------------- debug files
// The debug dump (shympi_debug, utils/shympi/mpi_debug.f90:114-146,367-415):
// per time block the record dtime, then per array three records --
// nn lmax type (1 integer, 2 real, 3 double), the name (80 chars), the array
// (lmax, nn) -- and the record 0 0 0 closing the block.  check_shympi_debug
// (utils/shympi/check_shympi_debug.f90) lists and compares such files.
struct shyhip_dbgfile {
    FILE *fp = nullptr;
    std::string err;
};

namespace {

int put_dbg(shyhip_dbgfile *d, std::initializer_list<Piece> parts) { return put_record(d->fp, d->err, parts); }

}  // namespace

extern "C" {

int shyhip_dbg_open(const char *path, shyhip_dbgfile **out) {
    if (!out || !path) return 1;
    auto *d = new shyhip_dbgfile;
    d->fp = std::fopen(path, "wb");
    if (!d->fp) {
        delete d;
        *out = nullptr;
        return 1;
    }
    *out = d;
    return 0;
}

int shyhip_dbg_time(shyhip_dbgfile *d, double dtime) {
    if (!d || !d->fp) return 1;
    return put_dbg(d, {{&dtime, sizeof dtime}});
}

// one array (lmax, nn), type 1 int32, 2 float, 3 double
int shyhip_dbg_record(shyhip_dbgfile *d, const char *name, int nn, int lmax, int type, const void *array) {
    if (!d || !d->fp || !array || nn <= 0 || lmax <= 0 || type < 1 || type > 3) return 1;
    const int head[3] = {nn, lmax, type};
    char t80[80];
    text80(t80, name);
    const size_t w = type == 3 ? 8 : 4;
    return put_dbg(d, {{head, sizeof head}}) || put_dbg(d, {{t80, 80}}) ||
           put_dbg(d, {{array, w * (size_t)nn * lmax}});
}

int shyhip_dbg_end_time(shyhip_dbgfile *d) {
    if (!d || !d->fp) return 1;
    const int z[3] = {0, 0, 0};
    return put_dbg(d, {{z, sizeof z}});
}

int shyhip_dbg_close(shyhip_dbgfile *d) {
    if (!d) return 0;
    const int rc = (d->fp && std::fclose(d->fp) != 0) ? 1 : 0;
    delete d;
    return rc;
}

// the records of the model's state that the reference's debug output
// writes after a step (shyfem_subs.f90:1210-1265, the arrays this path
// updates): znv, zenv, utlnv, vtlnv, wlnv, hdknv, hdenv, iwegv; the caller
// opens and closes the time block
int shyhip_dbg_write_state(shyhip_model *h, shyhip_dbgfile *d) {
    if (!h || !d) return 1;
    shy::Model &M = h->M;
    const int L = M.nlvdi, nk = M.nkn, ne = M.nel;
    std::vector<float> z(nk), ze(3 * (size_t)ne), u((size_t)L * ne), v((size_t)L * ne), w((size_t)(L + 1) * nk),
        hk((size_t)L * nk), he((size_t)L * ne);
    std::vector<int> iw(ne);
    if (shyhip_pull(h, SHYHIP_ZNV, z.data(), SHYHIP_F32, z.size()) ||
        shyhip_pull(h, SHYHIP_ZENV, ze.data(), SHYHIP_F32, ze.size()) ||
        shyhip_pull(h, SHYHIP_UTLNV, u.data(), SHYHIP_F32, u.size()) ||
        shyhip_pull(h, SHYHIP_VTLNV, v.data(), SHYHIP_F32, v.size()) ||
        shyhip_pull(h, SHYHIP_WLNV, w.data(), SHYHIP_F32, w.size()) ||
        shyhip_pull(h, SHYHIP_HDKNV, hk.data(), SHYHIP_F32, hk.size()) ||
        shyhip_pull(h, SHYHIP_HDENV, he.data(), SHYHIP_F32, he.size()) ||
        shyhip_pull(h, SHYHIP_IWEGV, iw.data(), SHYHIP_I32, iw.size()))
        return 1;
    return shyhip_dbg_record(d, "znv", nk, 1, 2, z.data()) ||
           shyhip_dbg_record(d, "zenv", ne, 3, 2, ze.data()) || shyhip_dbg_record(d, "utlnv", ne, L, 2, u.data()) ||
           shyhip_dbg_record(d, "vtlnv", ne, L, 2, v.data()) || shyhip_dbg_record(d, "wlnv", nk, L + 1, 2, w.data()) ||
           shyhip_dbg_record(d, "hdknv", nk, L, 2, hk.data()) || shyhip_dbg_record(d, "hdenv", ne, L, 2, he.data()) ||
           shyhip_dbg_record(d, "iwegv", ne, 1, 1, iw.data());
}

const char *shyhip_dbg_error(const shyhip_dbgfile *d) { return d ? d->err.c_str() : "null file"; }

}  // extern "C"
