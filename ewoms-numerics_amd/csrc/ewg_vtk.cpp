// ewg_vtk.cpp -- VTK output of a context's state as the reference writes it for the
// immiscible model (VtkMultiWriter, ewoms/numerics/io/vtkmultiwriter.hh:62-560, with the
// default fields of VtkMultiPhaseModule, io/vtkmultiphasemodule.hh:61-72, :311-337): one
// XML unstructured-grid file per call, "<simName>-<5-digit index>.vtu", with the cell
// fields pressure_w, pressure_n, density_w, density_n, saturation_w, saturation_n,
// relativePerm_w, relativePerm_n and porosity, and the collection "<simName>.pvd" listing
// every file written by the context with its time (vtkmultiwriter.hh:492-520).
// Values are written as ASCII Float64 with 17 significant digits (every double
// round-trips), so that fields can be compared cell by cell with a reference run.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ewg_capi.h"
#include "ewg_ctx.h"

namespace {

int fail(ewg_ctx *c, int code, const std::string &msg) {
    c->err = msg;
    return code;
}

void data_array(std::ostream &out, const char *name, const std::vector<double> &v) {
    out << "    <DataArray type=\"Float64\" Name=\"" << name
        << "\" NumberOfComponents=\"1\" format=\"ascii\">\n";
    for (size_t i = 0; i < v.size(); ++i) out << v[i] << ((i % 8 == 7) ? "\n" : " ");
    out << "\n    </DataArray>\n";
}

}  // namespace

extern "C" {

int ewg_write_vtk(ewg_ctx *c, const char *output_dir, const char *sim_name, double t) {
    if (!c || !output_dir || !sim_name) return EWG_INVALID_ARGUMENT;
    const ewg::Grid &g = c->g;
    const size_t N = g.N;
    std::vector<double> x(2 * N), phi(N), pe;
    std::vector<int32_t> region(N);
    if (int rc = ewg_get_solution(c, 0, x.data())) return rc;
    if (hipMemcpy(phi.data(), c->phi, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(region.data(), c->region, N * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(c, EWG_DEVICE_ERROR, "copying the static fields failed");
    if (c->pe) {
        pe.resize(N);
        if (hipMemcpy(pe.data(), c->pe, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(c, EWG_DEVICE_ERROR, "copying the entry pressures failed");
    }
    // ImmiscibleIntensiveQuantities::update (immiscibleintensivequantities.hh:87-145) per cell
    std::vector<double> pw(N), pn(N), rw(N), rn(N), sw(N), sn(N), kw(N), kn(N);
    for (size_t i = 0; i < N; ++i) {
        const ewg::MatReg &r = c->regs_host[c->nregions > 1 ? region[i] : 0];
        const double S = x[2 * i + 1];
        double krw, krn;
        ewg::relperm(r, S, krw, krn);
        pw[i] = x[2 * i];
        pn[i] = x[2 * i] + ewg::pcnw(r, pe.empty() ? 0.0 : pe[i], S);
        rw[i] = c->fluid.rho[0];
        rn[i] = c->fluid.rho[1];
        sw[i] = S;
        sn[i] = 1.0 - S;
        kw[i] = krw;
        kn[i] = krn;
    }
    std::string dir = output_dir;
    if (!dir.empty() && dir.back() != '/') dir += "/";
    std::ostringstream name;
    name << sim_name << "-" << std::setw(5) << std::setfill('0') << c->vtk_steps.size() << ".vtu";
    const std::string path = dir + name.str();
    std::ofstream out(path);
    if (!out.good()) return fail(c, EWG_INVALID_ARGUMENT, "cannot open " + path);
    out.precision(17);
    const int nx = g.nx, ny = g.ny, nz = g.dim == 3 ? g.nz : 1;
    const size_t npts = (size_t)(nx + 1) * (ny + 1) * (g.dim == 3 ? nz + 1 : 1);
    out << "<?xml version=\"1.0\"?>\n"
           "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
           " <UnstructuredGrid>\n";
    out << "  <Piece NumberOfCells=\"" << N << "\" NumberOfPoints=\"" << npts << "\">\n";
    out << "   <CellData Scalars=\"pressure_w\">\n";
    const std::pair<const char *, const std::vector<double> *> fields[] = {
        {"pressure_w", &pw},   {"pressure_n", &pn},   {"density_w", &rw},
        {"density_n", &rn},    {"saturation_w", &sw}, {"saturation_n", &sn},
        {"relativePerm_w", &kw}, {"relativePerm_n", &kn}, {"porosity", &phi}};
    for (const auto &f : fields) data_array(out, f.first, *f.second);
    out << "   </CellData>\n   <Points>\n"
           "    <DataArray type=\"Float64\" Name=\"Coordinates\" NumberOfComponents=\"3\" "
           "format=\"ascii\">\n";
    // YaspGrid vertices, lexicographic (x fastest)
    const int vz = g.dim == 3 ? nz + 1 : 1;
    for (int k = 0; k < vz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                out << i * g.h[0] << " " << j * g.h[1] << " " << (g.dim == 3 ? k * g.h[2] : 0.0)
                    << "\n";
    out << "    </DataArray>\n   </Points>\n   <Cells>\n"
           "    <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    auto vid = [&](int i, int j, int k) {
        return (long long)i + (long long)(nx + 1) * (j + (long long)(ny + 1) * k);
    };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                // VTK_QUAD / VTK_HEXAHEDRON corner order
                out << vid(i, j, k) << " " << vid(i + 1, j, k) << " " << vid(i + 1, j + 1, k)
                    << " " << vid(i, j + 1, k);
                if (g.dim == 3)
                    out << " " << vid(i, j, k + 1) << " " << vid(i + 1, j, k + 1) << " "
                        << vid(i + 1, j + 1, k + 1) << " " << vid(i, j + 1, k + 1);
                out << "\n";
            }
    const int corners = g.dim == 3 ? 8 : 4;
    out << "    </DataArray>\n"
           "    <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (size_t e = 0; e < N; ++e) out << (e + 1) * corners << ((e % 8 == 7) ? "\n" : " ");
    out << "\n    </DataArray>\n"
           "    <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (size_t e = 0; e < N; ++e) out << (g.dim == 3 ? 12 : 9) << ((e % 16 == 15) ? "\n" : " ");
    out << "\n    </DataArray>\n   </Cells>\n  </Piece>\n </UnstructuredGrid>\n</VTKFile>\n";
    out.close();
    if (!out) return fail(c, EWG_INVALID_ARGUMENT, "writing " + path + " failed");
    c->vtk_steps.emplace_back(t, name.str());
    // the collection of every file written so far (VtkMultiWriter's meta file)
    const std::string pvd = dir + sim_name + ".pvd";
    std::ofstream m(pvd);
    if (!m.good()) return fail(c, EWG_INVALID_ARGUMENT, "cannot open " + pvd);
    m.precision(16);
    m << "<?xml version=\"1.0\"?>\n"
         "<VTKFile type=\"Collection\"\n"
         "         version=\"0.1\"\n"
         "         byte_order=\"LittleEndian\"\n"
         "         compressor=\"vtkZLibDataCompressor\">\n"
         " <Collection>\n";
    for (const auto &st : c->vtk_steps)
        m << "   <DataSet timestep=\"" << st.first << "\" file=\"" << st.second << "\"/>\n";
    m << " </Collection>\n</VTKFile>\n";
    m.close();
    if (!m) return fail(c, EWG_INVALID_ARGUMENT, "writing " + pvd + " failed");
    return EWG_OK;
}

}  // extern "C"
