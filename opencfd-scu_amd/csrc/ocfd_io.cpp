// ocfd_io.cpp - the reference's input formats (host only).
//
//  * opencfd-scu.in: KEY = v1 v2 ... one key per line (src/parameters.c:
//    197-564, 667-792).  A line is a candidate only when it holds '=' and ends
//    with a newline; spaces are removed from the key; the last occurrence of a
//    key wins; '#'-prefixed keys never match; unknown keys are ignored;
//    FILTER_NPARA&/FILTER_RPARA& and ANA_EVENT&/ANA_NPARA&/ANA_RPARA& are
//    numbered families found by stripping digits from the key.
//  * Fortran-record 3-D arrays: per z plane, int32 nbytes | nx*ny f64 |
//    int32 nbytes (head/io_warp.h:6-20, src/OCFD_IO_mpi.c:562-740).
//  * Flow files: header record {int32 Istep, f64 tt} then d,u,v,w,T
//    (src/OCFD_IO.c:24-228).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ocfd.h"

namespace {

struct Parsed {
    std::map<std::string, std::string> kv;  // last occurrence wins
    std::vector<std::pair<std::string, std::string>> order;
};

std::string strip_spaces(const std::string& s) {
    std::string o;
    for (char c : s)
        if (c != ' ') o.push_back(c);
    return o;
}

bool parse_file(const char* path, Parsed& out) {
    FILE* fh = std::fopen(path, "r");
    if (!fh) return false;
    char buf[1000];
    while (std::fgets(buf, sizeof buf, fh)) {
        char* eq = std::strchr(buf, '=');
        char* lf = std::strchr(buf, '\n');
        if (!eq || !lf) continue;
        *lf = 0;
        std::string name = strip_spaces(std::string(buf, eq - buf));
        std::string value(eq + 1);
        out.kv[name] = value;
        out.order.push_back({name, value});
    }
    std::fclose(fh);
    return true;
}

std::vector<std::string> tokens(const std::string& s) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) j++;
        if (j > i) t.push_back(s.substr(i, j - i));
        i = j;
    }
    return t;
}

template <typename T>
void get_list(const Parsed& p, const char* key, T* dst, int n) {
    auto it = p.kv.find(key);
    if (it == p.kv.end()) return;
    auto t = tokens(it->second);
    for (int i = 0; i < n && i < (int)t.size(); i++) {
        if (std::is_same<T, int>::value) dst[i] = (T)std::atoi(t[i].c_str());
        else dst[i] = (T)std::atof(t[i].c_str());
    }
}

// src/parameters.c:567-637
int scheme_id(const std::string& s) {
    static const std::map<std::string, int> ids = {
        {"UP7", OCFD_UP7}, {"WENO5", OCFD_WENO5}, {"WENO7", OCFD_WENO7}, {"WENO7_SYMBO", OCFD_WENO7_SYMBO},
        {"WENO7_SYMBO_LIM", OCFD_WENO7_SYMBO_LIM}, {"NND2", OCFD_NND2}, {"OMP6_HR", OCFD_OMP6_HR},
        {"OMP6_LD", OCFD_OMP6_LD}, {"OMP6_CD8", OCFD_OMP6_CD8}, {"SCHEME_HYBRIDAUTO", OCFD_HYBRIDAUTO},
        {"CD6", OCFD_CD6}, {"CD8", OCFD_CD8}};
    auto it = ids.find(s);
    return it == ids.end() ? 0 : it->second;
}

bool is_digit_str(const std::string& s) {
    if (s.empty()) return false;
    for (char c : s)
        if (!std::isdigit((unsigned char)c)) return false;
    return true;
}

}  // namespace

extern "C" {

int ocfd_read_config(const char* path, ocfd_config* c) {
    if (!path || !c) return OCFD_ERR_ARG;
    Parsed p;
    if (!parse_file(path, p)) return OCFD_ERR_IO;
    std::memset(c, 0, sizeof(*c));
    {
        int v[3] = {0, 0, 0};
        get_list(p, "GRID_3D", v, 3);
        c->nx = v[0]; c->ny = v[1]; c->nz = v[2];
        int q[3] = {0, 0, 0};
        get_list(p, "PARALLEL_3D", q, 3);
        c->npx0 = q[0]; c->npy0 = q[1]; c->npz0 = q[2];
    }
    get_list(p, "STREAM", &c->stream_mode, 1);
    get_list(p, "TEST", &c->test, 1);
    get_list(p, "IPERIODIC", c->iperiodic, 3);
    get_list(p, "JAC_BOUND", c->jac_bound, 3);
    get_list(p, "DIF_BOUND", c->dif_bound, 6);
    get_list(p, "NON_REFLETION", c->non_ref, 6);
    {
        auto it = p.kv.find("SCHEME_INVIS");
        auto t = it == p.kv.end() ? std::vector<std::string>() : tokens(it->second);
        c->scheme_invis = t.empty() ? 0 : scheme_id(t[0]);
        it = p.kv.find("SCHEME_VIS");
        t = it == p.kv.end() ? std::vector<std::string>() : tokens(it->second);
        c->scheme_vis = t.empty() ? 0 : scheme_id(t[0]);
        // src/parameters.c:661-664: a wrong name is fatal
        if (c->scheme_invis < OCFD_UP7 || (c->scheme_vis != OCFD_CD6 && c->scheme_vis != OCFD_CD8))
            return OCFD_ERR_CONFIG;
    }
    if (c->scheme_invis == OCFD_HYBRIDAUTO) {
        get_list(p, "HY_DP_INTV", c->hy_dp_intv, 2);
        get_list(p, "HY_STYLE", &c->hy_style, 1);
        get_list(p, "HY_SMOOTH_DP", &c->hy_smooth_dp, 1);
        get_list(p, "HY_PATCH_ZONE", &c->hy_patch_zone, 1);
        if (c->hy_style != 1 && c->hy_style != 2) return OCFD_ERR_CONFIG;  // src/parameters.c:256-259
    }
    get_list(p, "RE", &c->re, 1);
    get_list(p, "AMA", &c->ama, 1);
    get_list(p, "GAMMA", &c->gamma, 1);
    get_list(p, "PR", &c->pr, 1);
    get_list(p, "T_REF", &c->t_ref, 1);
    get_list(p, "EPSL_SW", &c->epsl_sw, 1);
    get_list(p, "DT", &c->dt, 1);
    get_list(p, "END_TIME", &c->end_time, 1);
    get_list(p, "KSTEP_SHOW", &c->kstep_show, 1);
    get_list(p, "KSTEP_SAVE", &c->kstep_save, 1);
    get_list(p, "INIT_STAT", &c->init_stat, 1);
    get_list(p, "IBC", &c->ibc, 1);
    get_list(p, "CHARTERIC", &c->charteric, 1);
    {
        auto it = p.kv.find("BC_NPARA");
        if (it != p.kv.end()) {
            auto t = tokens(it->second);
            c->n_bc_npara = (int)std::min<size_t>(t.size(), 32);
            for (int i = 0; i < c->n_bc_npara; i++) c->bc_npara[i] = std::atoi(t[i].c_str());
        }
        it = p.kv.find("BC_RPARA");
        if (it != p.kv.end()) {
            auto t = tokens(it->second);
            c->n_bc_rpara = (int)std::min<size_t>(t.size(), 32);
            for (int i = 0; i < c->n_bc_rpara; i++) c->bc_rpara[i] = std::atof(t[i].c_str());
        }
    }
    if (c->t_ref <= 0) c->t_ref = 288.150;  // src/parameters.c:521
    // FILTER_NPARA& (src/parameters.c:320-343, ItemNUM 744-772): keys whose
    // name without digits is FILTER_NPARA, in file order.
    int nf = 0;
    for (auto& kv : p.order) {
        const std::string& k = kv.first;
        if (k.size() > 12 && k.compare(0, 12, "FILTER_NPARA") == 0 && is_digit_str(k.substr(12))) {
            if (nf >= OCFD_MAX_FILTERS) break;
            const std::string num = k.substr(12);
            auto t = tokens(p.kv["FILTER_NPARA" + num]);
            for (int q = 0; q < 11 && q < (int)t.size(); q++) c->filter_npara[nf][q] = std::atoi(t[q].c_str());
            auto r = tokens(p.kv["FILTER_RPARA" + num]);
            for (int q = 0; q < 3 && q < (int)r.size(); q++) c->filter_rpara[nf][q] = std::atof(r[q].c_str());
            nf++;
        }
    }
    c->nfilter = nf;
    // ANA_EVENT& = style Kstep, with ANA_NPARA& / ANA_RPARA& of the same
    // number (src/parameters.c:328-358): found like the filter family.
    int na = 0;
    for (auto& kv : p.order) {
        const std::string& k = kv.first;
        if (k.size() > 9 && k.compare(0, 9, "ANA_EVENT") == 0 && is_digit_str(k.substr(9))) {
            if (na >= OCFD_MAX_ANA) break;
            const std::string num = k.substr(9);
            auto e = tokens(p.kv["ANA_EVENT" + num]);
            for (int q = 0; q < 2 && q < (int)e.size(); q++) c->ana_event[na][q] = std::atoi(e[q].c_str());
            auto t = tokens(p.kv["ANA_NPARA" + num]);
            c->n_ana_npara[na] = (int)std::min<size_t>(t.size(), OCFD_ANA_NPARA);
            for (int q = 0; q < c->n_ana_npara[na]; q++) c->ana_npara[na][q] = std::atoi(t[q].c_str());
            auto r = tokens(p.kv["ANA_RPARA" + num]);
            c->n_ana_rpara[na] = (int)std::min<size_t>(r.size(), OCFD_ANA_NPARA);
            for (int q = 0; q < c->n_ana_rpara[na]; q++) c->ana_rpara[na][q] = std::atof(r[q].c_str());
            na++;
        }
    }
    c->nana = na;
    // extension key (ignored by the reference, which skips unknown keys)
    {
        auto it = p.kv.find("PRECISION");
        if (it != p.kv.end()) {
            auto t = tokens(it->second);
            if (!t.empty() && (t[0] == "MIXED" || t[0] == "1")) c->precision = OCFD_PREC_MIXED;
        }
    }
    return OCFD_OK;
}

int ocfd_read_records(const char* path, int nx, int ny, int nz, int nfield, double* out) {
    FILE* fh = std::fopen(path, "rb");
    if (!fh) return OCFD_ERR_IO;
    const size_t plane = (size_t)nx * ny;
    for (int f = 0; f < nfield; f++)
        for (int k = 0; k < nz; k++) {
            int32_t a = 0, b = 0;
            if (std::fread(&a, 4, 1, fh) != 1 ||
                std::fread(out + ((size_t)f * nz + k) * plane, sizeof(double), plane, fh) != plane ||
                std::fread(&b, 4, 1, fh) != 1) {
                std::fclose(fh);
                return OCFD_ERR_IO;
            }
        }
    std::fclose(fh);
    return OCFD_OK;
}

int ocfd_write_records(const char* path, int nx, int ny, int nz, int nfield, const double* in) {
    FILE* fh = std::fopen(path, "wb");
    if (!fh) return OCFD_ERR_IO;
    const size_t plane = (size_t)nx * ny;
    const int32_t nb = (int32_t)(plane * sizeof(double));
    for (int f = 0; f < nfield; f++)
        for (int k = 0; k < nz; k++) {
            std::fwrite(&nb, 4, 1, fh);
            std::fwrite(in + ((size_t)f * nz + k) * plane, sizeof(double), plane, fh);
            std::fwrite(&nb, 4, 1, fh);
        }
    std::fclose(fh);
    return OCFD_OK;
}

int ocfd_read_flow_file(const char* path, int nx, int ny, int nz, double* d, double* u, double* v, double* w,
                        double* T, int* istep, double* tt) {
    FILE* fh = std::fopen(path, "rb");
    if (!fh) return OCFD_ERR_IO;
    int32_t rec = 0, is = 0;
    double t = 0;
    if (std::fread(&rec, 4, 1, fh) != 1 || std::fread(&is, 4, 1, fh) != 1 || std::fread(&t, 8, 1, fh) != 1 ||
        std::fread(&rec, 4, 1, fh) != 1) {
        std::fclose(fh);
        return OCFD_ERR_IO;
    }
    double* outs[5] = {d, u, v, w, T};
    const size_t plane = (size_t)nx * ny;
    for (int f = 0; f < 5; f++)
        for (int k = 0; k < nz; k++) {
            int32_t a, b;
            if (std::fread(&a, 4, 1, fh) != 1 || std::fread(outs[f] + (size_t)k * plane, 8, plane, fh) != plane ||
                std::fread(&b, 4, 1, fh) != 1) {
                std::fclose(fh);
                return OCFD_ERR_IO;
            }
        }
    std::fclose(fh);
    if (istep) *istep = is;
    if (tt) *tt = t;
    return OCFD_OK;
}

int ocfd_write_flow_file(const char* path, int nx, int ny, int nz, const double* d, const double* u,
                         const double* v, const double* w, const double* T, int istep, double tt) {
    FILE* fh = std::fopen(path, "wb");
    if (!fh) return OCFD_ERR_IO;
    const int32_t hdr = 12;  // sizeof(int[3]) in OCFD_save, src/OCFD_IO.c:168-193
    const int32_t is = istep;
    std::fwrite(&hdr, 4, 1, fh);
    std::fwrite(&is, 4, 1, fh);
    std::fwrite(&tt, 8, 1, fh);
    std::fwrite(&hdr, 4, 1, fh);
    const double* ins[5] = {d, u, v, w, T};
    const size_t plane = (size_t)nx * ny;
    const int32_t nb = (int32_t)(plane * sizeof(double));
    for (int f = 0; f < 5; f++)
        for (int k = 0; k < nz; k++) {
            std::fwrite(&nb, 4, 1, fh);
            std::fwrite(ins[f] + (size_t)k * plane, 8, plane, fh);
            std::fwrite(&nb, 4, 1, fh);
        }
    std::fclose(fh);
    return OCFD_OK;
}

}  // extern "C"
