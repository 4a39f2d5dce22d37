// restart_io.hpp -- NekCEM's restart file, host side.
//
// The file cem_restart_out writes for a Maxwell run (src/io.F:411-490,
// through writeheader4 / writenodes4 / writecells4 / writefield4(_double)
// of src/io_co.c:115-554) and restart_swap reads back (src/io.F:637-788,
// src/io_co_read.c): a legacy-VTK binary unstructured grid, big-endian,
// of one rank's elements in global element order:
//
//   "# vtk DataFile Version 3.0 \n"
//   "<istep> <dumpno> <nx1> <np> <time %f> <dt %f> Electromagnetic Field  \n"
//   "BINARY \n" "DATASET UNSTRUCTURED_GRID \n"
//   "POINTS  <npts>  float  \n"  npts x (x, y, z) float32
//   "\nCELLS <ncell>  <9 ncell> \n"  ncell x (8, eight 0-based node ids) int32
//   "\nCELL_TYPES <ncell> \n"  ncell x 12 int32
//   "\nPOINT_DATA <npts> \n"
//   "VECTORS E   double \n"  npts x (Ex, Ey, Ez)   (or "float", float32)
//   " \n" "VECTORS H   double \n"  npts x (Hx, Hy, Hz)
//
// The cells are the (N)^3 GLL sub-hexahedra of each element in the node
// numbering of vtk_dump_header_swap (src/io_dumpvtk.F:608-686).  The time
// and dt of the header are printed with %f (six decimals), as the
// reference writes them; restart_swap adds the time read to the run's.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nekcem_hip.h"

namespace nkc {

struct RestartHeader {
    int32_t istep = 0, dumpno = 0, nx1 = 0, np = 1;
    double time = 0.0, dt = 0.0;
};

namespace rio {

inline void put_be(std::vector<char> &b, const void *p, size_t n) {
    const char *c = static_cast<const char *>(p);
    for (size_t i = 0; i < n; ++i) b.push_back(c[n - 1 - i]);
}

inline void put_str(std::vector<char> &b, const char *s) { b.insert(b.end(), s, s + std::strlen(s)); }

inline bool flush(FILE *f, std::vector<char> &b) {
    const bool ok = b.empty() || std::fwrite(b.data(), 1, b.size(), f) == b.size();
    b.clear();
    return ok;
}

template <class T>
inline T get_be(const unsigned char *p) {
    unsigned char c[sizeof(T)];
    for (size_t i = 0; i < sizeof(T); ++i) c[i] = p[sizeof(T) - 1 - i];
    T v;
    std::memcpy(&v, c, sizeof(T));
    return v;
}

// One header line (up to '\n') at pos, split on blanks.
inline bool line_tokens(const std::vector<unsigned char> &buf, size_t &pos, std::vector<std::string> &tok) {
    tok.clear();
    size_t e = pos;
    while (e < buf.size() && buf[e] != '\n') ++e;
    if (e >= buf.size()) return false;
    std::string cur;
    for (size_t i = pos; i < e; ++i) {
        if (buf[i] == ' ') {
            if (!cur.empty()) tok.push_back(cur);
            cur.clear();
        } else {
            cur.push_back((char)buf[i]);
        }
    }
    if (!cur.empty()) tok.push_back(cur);
    pos = e + 1;
    return true;
}

// Skip blanks and newlines between a data block and the next header line.
inline void skip_space(const std::vector<unsigned char> &buf, size_t &pos) {
    while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\n')) ++pos;
}

}  // namespace rio

// Write the restart file.  xyz: node coordinates, 3 components at stride
// xfs; E, H: the fields, 3 components at stride fs each (host memory).
inline int32_t restart_write(const char *path, const RestartHeader &hd, int n1, int64_t nelt, const double *xyz,
                             int64_t xfs, const double *E, const double *H, int64_t fs, bool dbl, std::string &err) {
    const int64_t n2 = (int64_t)n1 * n1, n3 = n2 * n1, npts = n3 * nelt;
    const int64_t ncell = (int64_t)(n1 - 1) * (n1 - 1) * (n1 - 1) * nelt;
    if (npts > INT32_MAX || 9 * ncell > INT32_MAX) {
        err = "mesh too large for the restart format's int32 counts";
        return NKC_ERR_ARG;
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("cannot open ") + path + " for writing";
        return NKC_ERR_ARG;
    }
    std::vector<char> b;
    b.reserve(1 << 22);
    bool ok = true;
    char line[512];
    rio::put_str(b, "# vtk DataFile Version 3.0 \n");
    std::snprintf(line, sizeof line, "%d %d %d %d %f %f Electromagnetic Field  \n", hd.istep, hd.dumpno, hd.nx1,
                  hd.np, hd.time, hd.dt);
    rio::put_str(b, line);
    rio::put_str(b, "BINARY \n");
    rio::put_str(b, "DATASET UNSTRUCTURED_GRID \n");
    std::snprintf(line, sizeof line, "POINTS  %lld  float  \n", (long long)npts);
    rio::put_str(b, line);
    for (int64_t p = 0; p < npts && ok; ++p) {
        for (int c = 0; c < 3; ++c) {
            const float v = (float)xyz[c * xfs + p];
            rio::put_be(b, &v, 4);
        }
        if (b.size() > (1 << 22)) ok = rio::flush(f, b);
    }
    std::snprintf(line, sizeof line, "\nCELLS %lld  %lld \n", (long long)ncell, (long long)(9 * ncell));
    rio::put_str(b, line);
    // vtk_dump_header_swap: per element and level, the (n1-1)^2 squares of
    // the level, skipping the last node of each row
    for (int64_t e = 0; e < nelt && ok; ++e) {
        for (int j = 1; j <= n1 - 1; ++j) {
            int64_t nf = (int64_t)(j - 1) * n2 + e * n3;
            for (int i = 1; i <= (n1 - 1) * (n1 - 1); ++i) {
                const int32_t c0[4] = {(int32_t)nf, (int32_t)(nf + 1), (int32_t)(nf + n1 + 1), (int32_t)(nf + n1)};
                const int32_t eight = 8;
                rio::put_be(b, &eight, 4);
                for (int q = 0; q < 4; ++q) rio::put_be(b, &c0[q], 4);
                for (int q = 0; q < 4; ++q) {
                    const int32_t v = c0[q] + (int32_t)n2;
                    rio::put_be(b, &v, 4);
                }
                nf += 1;
                if ((nf + 1) % n1 == 0) nf += 1;
            }
        }
        if (b.size() > (1 << 22)) ok = rio::flush(f, b);
    }
    std::snprintf(line, sizeof line, "\nCELL_TYPES %lld \n", (long long)ncell);
    rio::put_str(b, line);
    const int32_t hex = 12;
    for (int64_t c = 0; c < ncell && ok; ++c) {
        rio::put_be(b, &hex, 4);
        if (b.size() > (1 << 22)) ok = rio::flush(f, b);
    }
    std::snprintf(line, sizeof line, "\nPOINT_DATA %lld \n", (long long)npts);
    rio::put_str(b, line);
    const double *fld[2] = {E, H};
    const char *name[2] = {"E ", "H "};
    for (int k = 0; k < 2 && ok; ++k) {
        if (k) rio::put_str(b, " \n");
        std::snprintf(line, sizeof line, "VECTORS %s  %s \n", name[k], dbl ? "double" : "float");
        rio::put_str(b, line);
        for (int64_t p = 0; p < npts && ok; ++p) {
            for (int c = 0; c < 3; ++c) {
                const double v = fld[k][c * fs + p];
                if (dbl) {
                    rio::put_be(b, &v, 8);
                } else {
                    const float s = (float)v;
                    rio::put_be(b, &s, 4);
                }
            }
            if (b.size() > (1 << 22)) ok = rio::flush(f, b);
        }
    }
    // the writer leaves the last field's trailing " \n" unwritten
    ok = ok && rio::flush(f, b);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        err = std::string("write to ") + path + " failed";
        return NKC_ERR_ARG;
    }
    return NKC_OK;
}

// Read a restart file of this mesh: the header, then E and H into 3
// components at stride fs each (host memory).
inline int32_t restart_read(const char *path, int n1, int64_t nelt, RestartHeader &hd, double *E, double *H,
                            int64_t fs, std::string &err) {
    const int64_t n3 = (int64_t)n1 * n1 * n1, npts = n3 * nelt;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return NKC_ERR_ARG;
    }
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<unsigned char> buf(sz > 0 ? (size_t)sz : 0);
    const bool got = buf.empty() || std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    std::fclose(f);
    auto bad = [&](const std::string &m) {
        err = std::string(path) + ": " + m;
        return (int32_t)NKC_ERR_ARG;
    };
    if (!got) return bad("read failed");
    size_t pos = 0;
    std::vector<std::string> t;
    if (!rio::line_tokens(buf, pos, t) || t.size() < 2 || t[0] != "#" || t[1] != "vtk") return bad("not a VTK file");
    if (!rio::line_tokens(buf, pos, t) || t.size() < 6) return bad("no restart header line");
    hd.istep = std::atoi(t[0].c_str());
    hd.dumpno = std::atoi(t[1].c_str());
    hd.nx1 = std::atoi(t[2].c_str());
    hd.np = std::atoi(t[3].c_str());
    hd.time = std::atof(t[4].c_str());
    hd.dt = std::atof(t[5].c_str());
    if (!rio::line_tokens(buf, pos, t) || t.empty() || t[0] != "BINARY") return bad("not a binary VTK file");
    if (!rio::line_tokens(buf, pos, t) || t.size() < 2 || t[1] != "UNSTRUCTURED_GRID") return bad("not an unstructured grid");
    if (hd.nx1 != n1) return bad("written at nx1 = " + std::to_string(hd.nx1) + ", the mesh has " + std::to_string(n1));
    if (!rio::line_tokens(buf, pos, t) || t.size() < 3 || t[0] != "POINTS") return bad("no POINTS block");
    if (std::atoll(t[1].c_str()) != npts) return bad("POINTS " + t[1] + " != npts " + std::to_string(npts));
    const size_t pbytes = (t[2] == "double" ? 8 : 4) * 3 * (size_t)npts;
    if (pos + pbytes > buf.size()) return bad("truncated POINTS");
    pos += pbytes;
    rio::skip_space(buf, pos);
    if (!rio::line_tokens(buf, pos, t) || t.size() < 3 || t[0] != "CELLS") return bad("no CELLS block");
    const size_t cbytes = 4 * (size_t)std::atoll(t[2].c_str());
    if (pos + cbytes > buf.size()) return bad("truncated CELLS");
    pos += cbytes;
    rio::skip_space(buf, pos);
    if (!rio::line_tokens(buf, pos, t) || t.size() < 2 || t[0] != "CELL_TYPES") return bad("no CELL_TYPES block");
    const size_t tbytes = 4 * (size_t)std::atoll(t[1].c_str());
    if (pos + tbytes > buf.size()) return bad("truncated CELL_TYPES");
    pos += tbytes;
    rio::skip_space(buf, pos);
    if (!rio::line_tokens(buf, pos, t) || t.size() < 2 || t[0] != "POINT_DATA") return bad("no POINT_DATA");
    if (std::atoll(t[1].c_str()) != npts) return bad("POINT_DATA " + t[1] + " != npts");
    // restart_swap reads E into en, then H into hn
    double *dst[2] = {E, H};
    const char *want[2] = {"E", "H"};
    for (int k = 0; k < 2; ++k) {
        rio::skip_space(buf, pos);
        if (!rio::line_tokens(buf, pos, t) || t.size() < 3 || t[0] != "VECTORS" || t[1] != want[k])
            return bad(std::string("no VECTORS ") + want[k] + " block");
        const bool dbl = t[2] == "double";
        if (!dbl && t[2] != "float") return bad("VECTORS of type " + t[2]);
        const size_t w = dbl ? 8 : 4;
        if (pos + w * 3 * (size_t)npts > buf.size()) return bad(std::string("truncated VECTORS ") + want[k]);
        for (int64_t p = 0; p < npts; ++p)
            for (int c = 0; c < 3; ++c) {
                const unsigned char *q = &buf[pos + w * (3 * (size_t)p + c)];
                dst[k][c * fs + p] = dbl ? rio::get_be<double>(q) : (double)rio::get_be<float>(q);
            }
        pos += w * 3 * (size_t)npts;
    }
    return NKC_OK;
}

}  // namespace nkc
