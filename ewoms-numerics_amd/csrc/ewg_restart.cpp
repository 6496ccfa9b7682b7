// ewg_restart.cpp -- restart files of the reference's .ers format (Ewoms::Restart,
// ewoms/numerics/io/restart.hh:36-271) for a context: the simulator's clock, then the
// primary variables of every element (ecfvdiscretization.hh:186-202,
// fvbasediscretization.hh:1445-1482), so that a state can be handed to or taken from a
// run of the reference.
//
// File layout, as Restart::serializeBegin / serializeSectionBegin / serializeEntities
// write it (text, doubles with stream precision 20):
//   eWoms restart file: gridName='blubb' numCPUs=1 myRank=0 numElements=.. numEdges=..
//       numVertices=..                     (one line; the magic cookie)
//   <empty line>                            (end of the cookie section)
//   Simulator
//   episodeIdx episodeStartTime episodeLength startTime time timeStepIdx <space>
//   Entities: Codim 0
//   p_w S_w <space>                         (one line per element, element order)
//   <empty line>
// File name: <outputDir>/<simName>_time=<t>_rank=<rank>.ers with t in the stream's
// default format (restart.hh:70-85).
#include <cctype>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ewg_capi.h"
#include "ewg_ctx.h"

namespace {

int fail(ewg_ctx *c, int code, const std::string &msg) {
    c->err = msg;
    return code;
}

// Restart::magicRestartCookie_ (restart.hh:43-63) of a YaspGrid with the context's cells:
// numEdges is the number of entities of codimension dim - 1
std::string cookie(const ewg_ctx *c) {
    const long long nx = c->grid.nx, ny = c->grid.ny, nz = c->grid.nz;
    long long elements, edges, vertices;
    if (c->grid.dim == 2) {
        elements = nx * ny;
        edges = nx * (ny + 1) + (nx + 1) * ny;
        vertices = (nx + 1) * (ny + 1);
    } else {
        elements = nx * ny * nz;
        edges = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
        vertices = (nx + 1) * (ny + 1) * (nz + 1);
    }
    std::ostringstream oss;
    oss << "eWoms restart file: gridName='blubb' numCPUs=1 myRank=0 numElements=" << elements
        << " numEdges=" << edges << " numVertices=" << vertices;
    return oss.str();
}

// Restart::deserializeSectionEnd: the rest of the line must be white space
bool section_end(std::istream &in) {
    std::string rest;
    std::getline(in, rest);
    for (char ch : rest)
        if (!std::isspace(static_cast<unsigned char>(ch))) return false;
    return true;
}

}  // namespace

extern "C" {

int ewg_restart_file_name(const char *output_dir, const char *sim_name, double t, int32_t rank,
                          char *buf, int32_t len) {
    if (!output_dir || !sim_name || !buf || len < 1) return EWG_INVALID_ARGUMENT;
    std::string dir = output_dir;
    if (dir == ".") dir = "";
    else if (!dir.empty() && dir.back() != '/') dir += "/";
    std::ostringstream oss;
    oss << dir << sim_name << "_time=" << t << "_rank=" << rank << ".ers";
    const std::string name = oss.str();
    if ((int32_t)name.size() + 1 > len) return EWG_INVALID_ARGUMENT;
    std::copy(name.begin(), name.end(), buf);
    buf[name.size()] = '\0';
    return EWG_OK;
}

int ewg_write_restart(ewg_ctx *c, const char *path, const ewg_restart_clock *clock) {
    if (!c || !path || !clock) return EWG_INVALID_ARGUMENT;
    std::vector<double> x(2 * (size_t)c->g.N);
    if (int rc = ewg_get_solution(c, 0, x.data())) return rc;
    std::ofstream out(path);
    if (!out.good()) return fail(c, EWG_INVALID_ARGUMENT, std::string("cannot open ") + path);
    out.precision(20);
    out << cookie(c) << "\n" << "\n";
    out << "Simulator" << "\n";
    out << clock->episode_idx << " " << clock->episode_start_time << " "
        << clock->episode_length << " " << clock->start_time << " " << clock->time << " "
        << clock->time_step_idx << " ";
    out << "\n";
    out << "Entities: Codim 0" << "\n";
    for (size_t i = 0; i < (size_t)c->g.N; ++i) out << x[2 * i] << " " << x[2 * i + 1] << " " << "\n";
    out << "\n";
    out.close();
    if (!out) return fail(c, EWG_INVALID_ARGUMENT, std::string("writing ") + path + " failed");
    return EWG_OK;
}

int ewg_read_restart(ewg_ctx *c, const char *path, ewg_restart_clock *clock) {
    if (!c || !path || !clock) return EWG_INVALID_ARGUMENT;
    std::ifstream in(path);
    if (!in.good()) return fail(c, EWG_INVALID_ARGUMENT, std::string("Restart file '") + path +
                                                            "' could not be opened properly");
    in.seekg(0, std::ios::end);
    if (in.tellg() == 0) return fail(c, EWG_INVALID_ARGUMENT, std::string("Restart file '") + path + "' is empty");
    in.seekg(0, std::ios::beg);
    std::string line;
    std::getline(in, line);
    if (line != cookie(c))
        return fail(c, EWG_INVALID_ARGUMENT, "restart file of another grid: '" + line + "'");
    if (!section_end(in)) return fail(c, EWG_INVALID_ARGUMENT, "Encountered unread values while deserializing");
    std::getline(in, line);
    if (line != "Simulator") return fail(c, EWG_INVALID_ARGUMENT, "Could not start section 'Simulator'");
    ewg_restart_clock k{};
    in >> k.episode_idx >> k.episode_start_time >> k.episode_length >> k.start_time >> k.time >>
        k.time_step_idx;
    if (!in || !section_end(in))
        return fail(c, EWG_INVALID_ARGUMENT, "Could not read section 'Simulator'");
    std::getline(in, line);
    if (line != "Entities: Codim 0")
        return fail(c, EWG_INVALID_ARGUMENT, "Could not start section 'Entities: Codim 0'");
    std::vector<double> x(2 * (size_t)c->g.N);
    for (size_t i = 0; i < (size_t)c->g.N; ++i) {
        if (!in.good()) return fail(c, EWG_INVALID_ARGUMENT, "Restart file is corrupted");
        std::getline(in, line);
        std::istringstream ls(line);
        if (!(ls >> x[2 * i] >> x[2 * i + 1]))
            return fail(c, EWG_INVALID_ARGUMENT,
                        "Could not deserialize degree of freedom " + std::to_string(i));
    }
    if (!section_end(in)) return fail(c, EWG_INVALID_ARGUMENT, "Encountered unread values while deserializing");
    // EcfvDiscretization::deserialize: solution(1) = solution(0)
    if (int rc = ewg_set_solution(c, 0, x.data())) return rc;
    if (int rc = ewg_set_solution(c, 1, x.data())) return rc;
    *clock = k;
    return EWG_OK;
}

}  // extern "C"
