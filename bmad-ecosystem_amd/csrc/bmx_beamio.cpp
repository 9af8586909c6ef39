// Bmad ASCII beam files on the host (bmad/multiparticle/beam_file_io.f90): read the old
// "!ASCII::3" layout (read_beam_file, :380-804) and the column layout "ASCII::4"
// (read_beam_ascii, :834-1040), write the column layout (write_ascii_beam_file, :123-350).
// Files carry the phase-space coordinates, state and element index; the writer puts 17
// significant digits per coordinate so a checkpoint read back is bit-identical.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "bmx_internal.hpp"

using namespace bmx;

namespace {

// bmad_struct.f90:540-541 state_name(0:8), :403-404 location_name(0:3)
const char *const kStateName[9] = {"Pre_Born", "Alive",      "Lost",    "Lost_Neg_X", "Lost_Pos_X",
                                   "Lost_Neg_Y", "Lost_Pos_Y", "Lost_Pz", "Lost_Z"};
const char *const kLocationName[4] = {"GARBAGE!", "Upstream_End", "Downstream_End", "Inside"};

std::string upcase(std::string s)
{
    for (char &c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

int state_code(const std::string &name)
{
    const std::string u = upcase(name);
    for (int k = 0; k < 9; k++)
        if (upcase(kStateName[k]) == u) return k;
    char *end = nullptr;
    const long v = std::strtol(name.c_str(), &end, 10);
    return (end && *end == 0 && !name.empty()) ? (int)v : -1;
}

// Fortran-style reals: D exponents are accepted as E.
bool to_double(std::string t, double &v)
{
    for (char &c : t)
        if (c == 'd' || c == 'D') c = 'E';
    char *end = nullptr;
    errno = 0;
    v = std::strtod(t.c_str(), &end);
    return end && *end == 0 && !t.empty() && errno == 0;
}

std::vector<std::string> tokens(const std::string &line)
{
    std::istringstream is(line);
    std::vector<std::string> out;
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

// The text before a '!' comment.
std::string strip_comment(const std::string &line)
{
    const size_t k = line.find('!');
    return k == std::string::npos ? line : line.substr(0, k);
}

struct Particle6 {
    double v[6];
    int32_t state, ix_ele;
    double beta;
    bool has_beta;
};

struct Bunch {
    std::vector<Particle6> p;
    double charge_tot = 0;
};

// init_coord.f90:205-207 / convert_pc_to.f90:37-41
double beta_of(double pz, double p0c, double mc2)
{
    const double pc = p0c * (1 + pz);
    return pc / std::sqrt(pc * pc + mc2 * mc2);
}

// read_beam_file's old layout: header values with '!' comments, then per bunch BEGIN_BUNCH,
// species, charge_tot, z_center, t_center, one particle per line (vec(6), charge, state,
// spin(3), ix_ele, location), END_BUNCH.
int read_ascii3(std::istream &in, std::vector<Bunch> &bunches, std::string &err)
{
    std::string line;
    std::vector<std::string> head;
    while ((int)head.size() < 3 && std::getline(in, line)) {
        const std::string s = strip_comment(line);
        if (tokens(s).empty()) continue;
        head.push_back(s);
    }
    if (head.size() < 3) {
        err = "ASCII::3 header (ix_ele, n_bunch, n_particle) incomplete";
        return BMX_ERR_ARG;
    }
    const long n_bunch = std::strtol(head[1].c_str(), nullptr, 10);
    const long n_part = std::strtol(head[2].c_str(), nullptr, 10);
    if (n_bunch < 1 || n_part < 0) {
        err = "ASCII::3 header: bad n_bunch or n_particle";
        return BMX_ERR_ARG;
    }
    for (long b = 0; b < n_bunch; b++) {
        Bunch bunch;
        bool begun = false;
        while (std::getline(in, line))
            if (upcase(line).find("BEGIN_BUNCH") != std::string::npos) {
                begun = true;
                break;
            }
        if (!begun) {
            err = "no \"BEGIN_BUNCH\" marker found";
            return BMX_ERR_ARG;
        }
        std::vector<std::string> pre;  // species, charge_tot, z_center, t_center
        while ((int)pre.size() < 4 && std::getline(in, line)) {
            const std::string s = strip_comment(line);
            if (tokens(s).empty()) continue;
            pre.push_back(s);
        }
        if (pre.size() < 4 || !to_double(tokens(pre[1])[0], bunch.charge_tot)) {
            err = "bunch header (species, charge_tot, z_center, t_center) incomplete";
            return BMX_ERR_ARG;
        }
        while (std::getline(in, line)) {
            if (upcase(line).find("END_BUNCH") != std::string::npos) break;
            const std::vector<std::string> t = tokens(strip_comment(line));
            if (t.empty()) continue;
            if (t.size() < 8) {
                err = "particle line with fewer than 8 values: " + line;
                return BMX_ERR_ARG;
            }
            Particle6 p{};
            for (int k = 0; k < 6; k++)
                if (!to_double(t[k], p.v[k])) {
                    err = "bad number in: " + line;
                    return BMX_ERR_ARG;
                }
            p.state = (int32_t)std::strtol(t[7].c_str(), nullptr, 10);
            p.ix_ele = t.size() >= 12 ? (int32_t)std::strtol(t[11].c_str(), nullptr, 10) : 0;
            bunch.p.push_back(p);
        }
        bunches.push_back(std::move(bunch));
    }
    return BMX_OK;
}

// read_beam_ascii's column layout: '# name = value' parameters, a '#!' column-name line, then
// one particle per line; bunches start at '# ix_bunch' lines.
int read_ascii4(std::istream &in, std::vector<Bunch> &bunches, std::string &err)
{
    std::string line;
    std::vector<std::string> col;
    Particle6 p0{};
    p0.state = 1;
    double charge = 0;
    bool in_bunch = false;
    while (std::getline(in, line)) {
        std::string s = line;
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ')) s.pop_back();
        if (tokens(s).empty()) continue;
        if (s.compare(0, 2, "#!") == 0) {
            col = tokens(s.substr(2));
            continue;
        }
        if (s[0] == '#') {
            const size_t eq = s.find('=');
            if (eq == std::string::npos) continue;
            const std::vector<std::string> nm = tokens(s.substr(1, eq - 1));
            const std::vector<std::string> val = tokens(s.substr(eq + 1));
            if (nm.empty() || val.empty()) continue;
            const std::string name = nm[0];
            if (name == "ix_bunch") {
                bunches.emplace_back();
                in_bunch = true;
                p0 = Particle6{};
                p0.state = 1;
                charge = 0;
            } else if (name == "state") {
                p0.state = state_code(val[0]);
            } else if (name == "ix_ele") {
                p0.ix_ele = (int32_t)std::strtol(val[0].c_str(), nullptr, 10);
            } else if (name == "charge") {
                to_double(val[0], charge);
            } else if (name == "charge_tot") {
                if (in_bunch) to_double(val[0], bunches.back().charge_tot);
            }
            continue;
        }
        if (col.empty()) {
            err = "particle line before the '#!' column-name line";
            return BMX_ERR_ARG;
        }
        if (!in_bunch) {
            bunches.emplace_back();
            in_bunch = true;
        }
        const std::vector<std::string> t = tokens(s);
        if (t.size() < col.size()) {
            err = "particle line with fewer values than columns: " + line;
            return BMX_ERR_ARG;
        }
        Particle6 p = p0;
        double q = charge;
        static const char *const kVec[6] = {"x", "px", "y", "py", "z", "pz"};
        for (size_t c = 0; c < col.size(); c++) {
            const std::string &a = col[c];
            bool ok = true;
            int k = -1;
            for (int j = 0; j < 6; j++)
                if (a == kVec[j]) k = j;
            if (k >= 0) ok = to_double(t[c], p.v[k]);
            else if (a == "state") ok = (p.state = state_code(t[c])) >= 0;
            else if (a == "ix_ele") p.ix_ele = (int32_t)std::strtol(t[c].c_str(), nullptr, 10);
            else if (a == "beta") ok = p.has_beta = to_double(t[c], p.beta);
            else if (a == "charge") ok = to_double(t[c], q);
            if (!ok) {
                err = "bad value in column " + a + ": " + line;
                return BMX_ERR_ARG;
            }
        }
        bunches.back().p.push_back(p);
        bunches.back().charge_tot += q;
    }
    return BMX_OK;
}

int read_file(const char *path, std::vector<Bunch> &bunches, std::string &err)
{
    std::ifstream in(path);
    if (!in) {
        err = std::string("cannot open beam file: ") + path;
        return BMX_ERR_ARG;
    }
    std::string first;
    std::streampos start = in.tellg();
    while (std::getline(in, first))
        if (!tokens(first).empty()) break;
    if (upcase(first).find("!ASCII::3") == 0) return read_ascii3(in, bunches, err);
    in.clear();
    in.seekg(start);
    return read_ascii4(in, bunches, err);
}

}  // namespace

extern "C" {

int32_t bmx_beam_file_info(const char *path, int32_t *n_bunch, int64_t *n_particle)
{
    if (!path) return set_error(BMX_ERR_ARG, "bmx_beam_file_info: null path");
    std::vector<Bunch> b;
    std::string err;
    const int rc = read_file(path, b, err);
    if (rc != BMX_OK) return set_error(rc, "bmx_beam_file_info: " + err);
    if (n_bunch) *n_bunch = (int32_t)b.size();
    if (n_particle) *n_particle = b.empty() ? 0 : (int64_t)b[0].p.size();
    return BMX_OK;
}

int32_t bmx_beam_file_read(const char *path, int32_t ix_bunch, const bmx_beam_view *host, int64_t cap,
                           double p0c, double mc2, int64_t *n_out, double *charge_tot)
{
    if (!path || !host || cap < 0 || ix_bunch < 1 || !(p0c > 0) || !(mc2 > 0))
        return set_error(BMX_ERR_ARG, "bmx_beam_file_read: bad argument");
    std::vector<Bunch> b;
    std::string err;
    const int rc = read_file(path, b, err);
    if (rc != BMX_OK) return set_error(rc, "bmx_beam_file_read: " + err);
    if (ix_bunch > (int32_t)b.size()) return set_error(BMX_ERR_ARG, "bmx_beam_file_read: no such bunch");
    const Bunch &bu = b[ix_bunch - 1];
    const int64_t n = (int64_t)bu.p.size();
    if (n > cap) return set_error(BMX_ERR_ARG, "bmx_beam_file_read: more particles than cap");
    double *v[6] = {(double *)host->x, (double *)host->px, (double *)host->y,
                    (double *)host->py, (double *)host->z, (double *)host->pz};
    for (int64_t i = 0; i < n; i++) {
        const Particle6 &p = bu.p[i];
        for (int k = 0; k < 6; k++) v[k][i] = p.v[k];
        ((double *)host->beta)[i] = p.has_beta ? p.beta : beta_of(p.v[5], p0c, mc2);
        ((int32_t *)host->state)[i] = p.state;
        // a lost particle's ix_ele is the element it was lost at
        ((int32_t *)host->loss_ele)[i] = p.state != BMX_STATE_ALIVE ? p.ix_ele : 0;
        ((int32_t *)host->loss_turn)[i] = 0;
    }
    if (n_out) *n_out = n;
    if (charge_tot) *charge_tot = bu.charge_tot;
    return BMX_OK;
}

int32_t bmx_beam_file_write(const char *path, const bmx_beam_view *host, int64_t n, const char *species,
                            double charge, double p0c, int32_t ix_ele, int32_t append)
{
    if (!path || !host || n < 0 || !species)
        return set_error(BMX_ERR_ARG, "bmx_beam_file_write: bad argument");
    FILE *f = std::fopen(path, append ? "a" : "w");
    if (!f) return set_error(BMX_ERR_ARG, std::string("bmx_beam_file_write: cannot open ") + path);
    const double *v[6] = {(const double *)host->x, (const double *)host->px, (const double *)host->y,
                          (const double *)host->py, (const double *)host->z, (const double *)host->pz};
    const int32_t *st = (const int32_t *)host->state;
    const int32_t *le = (const int32_t *)host->loss_ele;
    int64_t n_alive = 0;
    for (int64_t i = 0; i < n; i++) n_alive += st[i] == BMX_STATE_ALIVE;
    // header as write_ascii_beam_file.f90 writes it (:200-213), constant columns as parameters
    std::fprintf(f, "# ix_bunch       =%6d\n", 1);
    std::fprintf(f, "# n_particle     =%6lld\n", (long long)n);
    std::fprintf(f, "# n_alive        =%6lld\n", (long long)n_alive);
    std::fprintf(f, "# n_dead         =%6lld\n", (long long)(n - n_alive));
    std::fprintf(f, "# %-15s=%24.16E\n", "charge", charge);
    std::fprintf(f, "# %-15s= %d\n", "ix_branch", 0);
    std::fprintf(f, "# %-15s= %d\n", "direction", 1);
    std::fprintf(f, "# %-15s= %d\n", "time_dir", 1);
    std::fprintf(f, "# %-15s= %s\n", "species", species);
    std::fprintf(f, "# %-15s= %s\n", "location", kLocationName[2]);
    std::fprintf(f, "#!%6s%24s%24s%24s%24s%24s%24s%24s%24s%14s%8s\n", "index", "x", "px", "y", "py", "z",
                 "pz", "p0c", "beta", "state", "ix_ele");
    for (int64_t i = 0; i < n; i++) {
        const int s = st[i];
        const char *sn = (s >= 0 && s <= 8) ? kStateName[s] : "Lost";
        std::fprintf(f, "%8lld", (long long)(i + 1));
        for (int k = 0; k < 6; k++) std::fprintf(f, "%24.16E", v[k][i]);
        std::fprintf(f, "%24.16E%24.16E%14s%8d\n", p0c, ((const double *)host->beta)[i], sn,
                     s == BMX_STATE_ALIVE ? ix_ele : le[i]);
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    if (!ok) return set_error(BMX_ERR_ARG, std::string("bmx_beam_file_write: write error on ") + path);
    return BMX_OK;
}

}  // extern "C"
