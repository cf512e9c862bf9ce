// Output files and clump-file readers of the DEMSolver facade: the CSV sphere, clump and contact writers of the
// reference's dynamic thread (src/DEM/dT.cpp:896-1180, dispatched by APIPublic.cpp:599-655) and its
// ReadClumpXyzFromCsv / ReadClumpQuatFromCsv (src/DEM/API.h:326-371). Numbers are written as std::ostream writes a
// float by default (6 significant digits) from the same float arithmetic, so a file matches the reference's text for
// the same state.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <limits>
#include <sstream>

#include "DEMSolver.h"
#include "host_math.h"

namespace deme {

namespace {

using hostmath::host_rotate;

void put(std::string& s, float v) {
    char b[32];
    const int n = std::snprintf(b, sizeof(b), "%g", (double)v);
    s.append(b, (size_t)n);
}
void put3(std::string& s, float x, float y, float z) {
    s += ',';
    put(s, x);
    s += ',';
    put(s, y);
    s += ',';
    put(s, z);
}

// Owner centre of mass as the writers compute it: hostVoxelIDToPosition<float> (HostSideHelpers.hpp:307-324) with the
// voxel size and sub-voxel unit rounded to float, then the domain corner added in float (dT.cpp:940-946).
float3 com_float(const deme_sim_params& p, uint64_t id, uint16_t sx, uint16_t sy, uint16_t sz) {
    const uint64_t vx = id & (((uint64_t)1 << p.nvXp2) - 1);
    const uint64_t vy = (id >> p.nvXp2) & (((uint64_t)1 << p.nvYp2) - 1);
    const uint64_t vz = id >> (p.nvXp2 + p.nvYp2);
    const float vs = (float)p.voxelSize_exact, l = (float)p.l_exact;
    const float X = (float)vx * vs + (float)sx * l;
    const float Y = (float)vy * vs + (float)sy * l;
    const float Z = (float)vz * vs + (float)sz * l;
    return make_float3(X + p.LBF[0], Y + p.LBF[1], Z + p.LBF[2]);
}

void write_text(const std::string& name, const std::string& text, bool binary) {
    std::ofstream f(name, binary ? (std::ios::out | std::ios::binary) : std::ios::out);
    if (!f) throw std::runtime_error("Cannot open " + name + " for writing.");
    f << text;
    if (!f) throw std::runtime_error("Writing " + name + " failed.");
}

// ---- reading ----------------------------------------------------------------------------------------------------
// The reader behaves as the reference's io::CSVReader<N, trim_chars<' ', '\t'>, no_quote_escape<','>,
// throw_on_overflow, empty_line_comment> with read_header(ignore_extra_column, ...) (src/core/utils/csv.hpp, a
// modified copy of Ben Strasser's fast-cpp-csv-parser): lines end at '\n' with one trailing '\r' dropped; lines of
// only spaces and tabs are skipped; fields are split at every comma and trimmed of spaces and tabs; the header is the
// first non-blank line.

struct CsvError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const std::string& file, unsigned line, const std::string& what) {
    std::ostringstream ss;
    ss << "Reading " << file;
    if (line) ss << " at line " << line;
    ss << ": " << what;
    throw CsvError(ss.str());
}

bool blank_line(const std::string& l) {
    for (char ch : l)
        if (ch != ' ' && ch != '\t') return false;
    return true;
}

void split_fields(const std::string& l, std::vector<std::string>& out) {
    out.clear();
    size_t b = 0;
    while (true) {
        const size_t e = l.find(',', b);
        size_t s = b, t = e == std::string::npos ? l.size() : e;
        while (s < t && (l[s] == ' ' || l[s] == '\t')) s++;
        while (t > s && (l[t - 1] == ' ' || l[t - 1] == '\t')) t--;
        out.emplace_back(l, s, t - s);
        if (e == std::string::npos) break;
        b = e + 1;
    }
}

// Decimal text to float as the reference's reader converts it (parse_float<float>, csv.hpp:937-997): integer digits
// accumulate in float as x = 10 x + d; the k-th fraction digit adds d * p_k with p_k = p_{k-1} / 10 in float; an
// exponent e (saturating at the int range) multiplies by 10^e, or 0.1^|e| for negative e, the power formed by binary
// powering in float. This is not a correctly rounded conversion: it can differ from strtof in the last bits, and
// equals the reference's value for the same text. An empty field reads as 0. Returns false for text that is not a
// number. The reference's loop never ends for the lowest int exponent; here that exponent is applied like any other.
bool csv_float(const char* c, float& out) {
    bool neg = false;
    if (*c == '-') {
        neg = true;
        ++c;
    } else if (*c == '+') {
        ++c;
    }
    float x = 0;
    while ('0' <= *c && *c <= '9') {
        const int d = *c - '0';
        x *= 10;
        x += d;
        ++c;
    }
    if (*c == '.' || *c == ',') {
        ++c;
        float p = 1;
        while ('0' <= *c && *c <= '9') {
            p /= 10;
            const int d = *c - '0';
            ++c;
            x += d * p;
        }
    }
    if (*c == 'e' || *c == 'E') {
        ++c;
        // signed integer with saturation (parse_signed_integer<set_to_max_on_overflow>, csv.hpp:890-912): the digits
        // after a saturation are not examined
        int64_t e = 0;
        bool sat = false;
        if (*c == '-') {
            ++c;
            for (; *c && !sat; ++c) {
                if (*c < '0' || *c > '9') return false;
                const int d = *c - '0';
                if (e < ((int64_t)std::numeric_limits<int>::min() + d) / 10) {
                    e = std::numeric_limits<int>::min();
                    sat = true;
                } else {
                    e = 10 * e - d;
                }
            }
        } else {
            if (*c == '+') ++c;
            for (; *c && !sat; ++c) {
                if (*c < '0' || *c > '9') return false;
                const int d = *c - '0';
                if (e > ((int64_t)std::numeric_limits<int>::max() - d) / 10) {
                    e = std::numeric_limits<int>::max();
                    sat = true;
                } else {
                    e = 10 * e + d;
                }
            }
        }
        if (e != 0) {
            float base = e < 0 ? 0.1f : 10.0f;
            uint64_t m = e < 0 ? (uint64_t)(-e) : (uint64_t)e;
            while (m != 1) {
                if ((m & 1) == 0) {
                    base = base * base;
                    m >>= 1;
                } else {
                    x *= base;
                    --m;
                }
            }
            x *= base;
        }
    } else if (*c != '\0') {
        return false;
    }
    out = neg ? -x : x;
    return true;
}

// Reads the named columns of a CSV file: one string column, then nf float columns, row by row.
template <size_t NF>
void read_columns(const std::string& file, const std::string& strName, const std::array<std::string, NF>& names,
                  const std::function<void(const std::string&, const std::array<float, NF>&)>& row) {
    std::ifstream in(file, std::ios::in | std::ios::binary);
    if (!in) fail(file, 0, "cannot open the file");
    std::string l;
    unsigned lineNo = 0;
    auto next = [&](std::string& s) {
        while (std::getline(in, s)) {
            lineNo++;
            if (!s.empty() && s.back() == '\r') s.pop_back();
            if (!blank_line(s)) return true;
        }
        return false;
    };
    if (!next(l)) fail(file, 0, "the header line is missing");
    std::vector<std::string> f;
    split_fields(l, f);
    // colOf[k]: the field index of requested column k (0 = the string column, then the floats)
    std::array<int, NF + 1> colOf;
    colOf.fill(-1);
    for (size_t i = 0; i < f.size(); i++) {
        for (size_t k = 0; k <= NF; k++) {
            const std::string& want = k == 0 ? strName : names[k - 1];
            if (f[i] != want) continue;
            if (colOf[k] != -1) fail(file, lineNo, "column '" + want + "' appears twice in the header");
            colOf[k] = (int)i;
            break;
        }
    }
    for (size_t k = 0; k <= NF; k++)
        if (colOf[k] == -1) fail(file, lineNo, "column '" + (k == 0 ? strName : names[k - 1]) + "' is missing");
    const size_t nCols = f.size();
    std::array<float, NF> v;
    while (next(l)) {
        split_fields(l, f);
        if (f.size() < nCols) fail(file, lineNo, "too few columns");
        if (f.size() > nCols) fail(file, lineNo, "too many columns");
        for (size_t k = 0; k < NF; k++) {
            const std::string& t = f[colOf[k + 1]];
            if (!csv_float(t.c_str(), v[k]))
                fail(file, lineNo, "the value '" + t + "' of column '" + names[k] + "' is not a number");
        }
        row(f[colOf[0]], v);
    }
}

}  // namespace

// ---- writers ----------------------------------------------------------------------------------------------------

void DEMSolver::writeSpheresCsv(std::string& s) {
    syncHostState();
    s = OUTPUT_FILE_X_COL_NAME + "," + OUTPUT_FILE_Y_COL_NAME + "," + OUTPUT_FILE_Z_COL_NAME + "," +
        OUTPUT_FILE_R_COL_NAME;
    if (m_out_content & OUTPUT_CONTENT::ABSV) s += ",absv";
    if (m_out_content & OUTPUT_CONTENT::VEL) s += ",v_x,v_y,v_z";
    if (m_out_content & OUTPUT_CONTENT::FAMILY) s += ",family";
    s += '\n';
    s.reserve(s.size() + nSpheresGM * 96);
    for (size_t i = 0; i < nSpheresGM; i++) {
        const uint32_t o = m_sph_owner[i];
        const unsigned int fam = h_fam[o];
        if (m_no_output_families.count(fam)) continue;
        const float3 c = com_float(m_params, h_voxel[o], h_lx[o], h_ly[o], h_lz[o]);
        float3 d = m_sph_rel[i];
        host_rotate(d.x, d.y, d.z, h_qw[o], h_qx[o], h_qy[o], h_qz[o]);
        put(s, c.x + d.x);
        s += ',';
        put(s, c.y + d.y);
        s += ',';
        put(s, c.z + d.z);
        s += ',';
        put(s, m_sph_r[i]);
        const float vx = h_vx[o], vy = h_vy[o], vz = h_vz[o];
        if (m_out_content & OUTPUT_CONTENT::ABSV) {
            s += ',';
            put(s, std::sqrt(vx * vx + vy * vy + vz * vz));
        }
        if (m_out_content & OUTPUT_CONTENT::VEL) put3(s, vx, vy, vz);
        if (m_out_content & OUTPUT_CONTENT::FAMILY) s += "," + std::to_string(fam);
        s += '\n';
    }
}

void DEMSolver::writeClumpsCsv(std::string& s) {
    syncHostState();
    s = OUTPUT_FILE_X_COL_NAME + "," + OUTPUT_FILE_Y_COL_NAME + "," + OUTPUT_FILE_Z_COL_NAME + ",Qw,Qx,Qy,Qz," +
        OUTPUT_FILE_CLUMP_TYPE_NAME;
    if (m_out_content & OUTPUT_CONTENT::ABSV) s += ",absv";
    if (m_out_content & OUTPUT_CONTENT::VEL) s += ",v_x,v_y,v_z";
    if (m_out_content & OUTPUT_CONTENT::FAMILY) s += ",family";
    s += '\n';
    s.reserve(s.size() + nOwnerClumps * 128);
    for (size_t o = 0; o < nOwnerClumps; o++) {  // clumps are the first owners; external objects follow
        const unsigned int fam = h_fam[o];
        if (m_no_output_families.count(fam)) continue;
        const float3 c = com_float(m_params, h_voxel[o], h_lx[o], h_ly[o], h_lz[o]);
        put(s, c.x);
        s += ',';
        put(s, c.y);
        s += ',';
        put(s, c.z);
        s += ',';
        put(s, h_qw[o]);
        put3(s, h_qx[o], h_qy[o], h_qz[o]);
        s += ',';
        s += m_templates.at(m_clump_template[o])->m_name;
        const float vx = h_vx[o], vy = h_vy[o], vz = h_vz[o];
        if (m_out_content & OUTPUT_CONTENT::ABSV) {
            s += ',';
            put(s, std::sqrt(vx * vx + vy * vy + vz * vz));
        }
        if (m_out_content & OUTPUT_CONTENT::VEL) put3(s, vx, vy, vz);
        if (m_out_content & OUTPUT_CONTENT::FAMILY) s += "," + std::to_string(fam);
        s += '\n';
    }
}

void DEMSolver::writeContactsCsv(std::string& s) {
    syncHostState();
    const unsigned int cf = m_cnt_out_content;
    s = OUTPUT_FILE_OWNER_1_NAME + "," + OUTPUT_FILE_OWNER_2_NAME + "," + OUTPUT_FILE_CNT_TYPE_NAME;
    if (cf & CNT_OUTPUT_CONTENT::FORCE)
        s += "," + OUTPUT_FILE_FORCE_X_NAME + "," + OUTPUT_FILE_FORCE_Y_NAME + "," + OUTPUT_FILE_FORCE_Z_NAME;
    if (cf & CNT_OUTPUT_CONTENT::POINT)
        s += "," + OUTPUT_FILE_X_COL_NAME + "," + OUTPUT_FILE_Y_COL_NAME + "," + OUTPUT_FILE_Z_COL_NAME;
    if (cf & CNT_OUTPUT_CONTENT::NORMAL)
        s += "," + OUTPUT_FILE_NORMAL_X_NAME + "," + OUTPUT_FILE_NORMAL_Y_NAME + "," + OUTPUT_FILE_NORMAL_Z_NAME;
    if (cf & CNT_OUTPUT_CONTENT::TORQUE_ONLY_FORCE)
        s += "," + OUTPUT_FILE_TOF_X_NAME + "," + OUTPUT_FILE_TOF_Y_NAME + "," + OUTPUT_FILE_TOF_Z_NAME;
    s += '\n';

    deme_counts cnt;
    check(deme_get_counts(m_ctx, &cnt), "deme_get_counts");
    const size_t C = cnt.nSteps ? cnt.nContacts : 0;  // no contact list before the first step
    if (C == 0) return;
    std::vector<uint32_t> idA(C), idB(C);
    std::vector<uint8_t> type(C);
    check(deme_download_contacts(m_ctx, idA.data(), idB.data(), type.data(), nullptr), "deme_download_contacts");
    const bool needRec = cf & (CNT_OUTPUT_CONTENT::FORCE | CNT_OUTPUT_CONTENT::POINT | CNT_OUTPUT_CONTENT::NORMAL |
                               CNT_OUTPUT_CONTENT::TORQUE_ONLY_FORCE);
    std::vector<float> F, T, cpA;
    if (needRec) {
        F.resize(3 * C);
        T.resize(3 * C);
        cpA.resize(3 * C);
        check(deme_download_forces(m_ctx, F.data(), T.data(), cpA.data(), nullptr), "deme_download_forces");
    }
    s.reserve(s.size() + C * 160);
    for (size_t i = 0; i < C; i++) {
        const uint8_t t = type[i];
        if (t == DEME_NOT_A_CONTACT) continue;
        const uint32_t oA = m_sph_owner.at(idA[i]);
        uint32_t oB;
        const char* name;
        if (t == DEME_SPHERE_SPHERE_CONTACT) {
            oB = m_sph_owner.at(idB[i]);
            name = "SS";
        } else if (t == DEME_SPHERE_MESH_CONTACT) {
            throw std::runtime_error("Sphere-mesh contacts do not occur on this path.");
        } else {
            oB = m_anal_owner.at(idB[i]);
            name = "SA";
        }
        s += std::to_string(toUser(oA));
        s += ',';
        s += std::to_string(toUser(oB));
        s += ',';
        s += name;
        if (cf & CNT_OUTPUT_CONTENT::FORCE) put3(s, F[3 * i], F[3 * i + 1], F[3 * i + 2]);
        if (cf & (CNT_OUTPUT_CONTENT::POINT | CNT_OUTPUT_CONTENT::NORMAL)) {
            // the contact point is kept in A's owner frame: rotate by A's orientation and add A's centre of mass
            // (the reference indexes the orientation by the contact number there, dT.cpp:1131-1134; owner A's is
            // the one that places the point)
            const float qw = h_qw[oA], qx = h_qx[oA], qy = h_qy[oA], qz = h_qz[oA];
            const float3 c = com_float(m_params, h_voxel[oA], h_lx[oA], h_ly[oA], h_lz[oA]);
            float px = cpA[3 * i], py = cpA[3 * i + 1], pz = cpA[3 * i + 2];
            host_rotate(px, py, pz, qw, qx, qy, qz);
            px += c.x;
            py += c.y;
            pz += c.z;
            if (cf & CNT_OUTPUT_CONTENT::POINT) put3(s, px, py, pz);
            if (cf & CNT_OUTPUT_CONTENT::NORMAL) {
                float3 d = m_sph_rel[idA[i]];
                host_rotate(d.x, d.y, d.z, qw, qx, qy, qz);
                const float nx = px - (c.x + d.x), ny = py - (c.y + d.y), nz = pz - (c.z + d.z);
                const float dd = nx * nx + ny * ny + nz * nz;
                // a pair of the list that does not touch has no contact point; its normal is written as zero
                if (dd > 0.f) {
                    const float inv = 1.0f / std::sqrt(dd);
                    put3(s, nx * inv, ny * inv, nz * inv);
                } else {
                    put3(s, 0.f, 0.f, 0.f);
                }
            }
        }
        if (cf & CNT_OUTPUT_CONTENT::TORQUE_ONLY_FORCE) put3(s, T[3 * i], T[3 * i + 1], T[3 * i + 2]);
        s += '\n';
    }
}

// The writers refresh the host copy of the owner state, a cache, so they stay const as in the reference.
void DEMSolver::WriteSphereFile(const std::string& outfilename) const {
    std::string s;
    switch (m_out_format) {
        case OUTPUT_FORMAT::CSV:
            const_cast<DEMSolver*>(this)->writeSpheresCsv(s);
            write_text(outfilename, s, false);
            break;
        case OUTPUT_FORMAT::BINARY:
            write_text(outfilename, "", true);  // not implemented by the reference either: an empty file
            break;
        default:
            throw std::runtime_error("The CHPF output format is not available; use OUTPUT_FORMAT::CSV.");
    }
}

void DEMSolver::WriteClumpFile(const std::string& outfilename) const {
    std::string s;
    switch (m_out_format) {
        case OUTPUT_FORMAT::CSV:
            const_cast<DEMSolver*>(this)->writeClumpsCsv(s);
            write_text(outfilename, s, false);
            break;
        case OUTPUT_FORMAT::BINARY:
            write_text(outfilename, "", true);
            break;
        default:
            throw std::runtime_error("The CHPF output format is not available; use OUTPUT_FORMAT::CSV.");
    }
}

void DEMSolver::WriteContactFile(const std::string& outfilename) const {
    if (m_cnt_out_format != OUTPUT_FORMAT::CSV)
        throw std::runtime_error(
            "Contact pair output file format is unknown or not implemented. Please re-set it via SetOutputFormat.");
    std::string s;
    const_cast<DEMSolver*>(this)->writeContactsCsv(s);
    write_text(outfilename, s, false);
}

// ---- readers ----------------------------------------------------------------------------------------------------

std::unordered_map<std::string, std::vector<float3>> DEMSolver::ReadClumpXyzFromCsv(const std::string& infilename,
                                                                                   const std::string& clump_header,
                                                                                   const std::string& x_header,
                                                                                   const std::string& y_header,
                                                                                   const std::string& z_header) {
    std::unordered_map<std::string, std::vector<float3>> m;
    read_columns<3>(infilename, clump_header, {x_header, y_header, z_header},
                    [&](const std::string& t, const std::array<float, 3>& v) {
                        m[t].push_back(make_float3(v[0], v[1], v[2]));
                    });
    return m;
}

std::unordered_map<std::string, std::vector<float4>> DEMSolver::ReadClumpQuatFromCsv(const std::string& infilename,
                                                                                    const std::string& clump_header,
                                                                                    const std::string& qw_header,
                                                                                    const std::string& qx_header,
                                                                                    const std::string& qy_header,
                                                                                    const std::string& qz_header) {
    std::unordered_map<std::string, std::vector<float4>> m;
    read_columns<4>(infilename, clump_header, {qw_header, qx_header, qy_header, qz_header},
                    [&](const std::string& t, const std::array<float, 4>& v) {
                        float4 q;
                        q.w = v[0];
                        q.x = v[1];
                        q.y = v[2];
                        q.z = v[3];
                        m[t].push_back(q);
                    });
    return m;
}

}  // namespace deme
