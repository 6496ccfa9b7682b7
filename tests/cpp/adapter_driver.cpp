// adapter_driver.cpp -- drives the C++ drop-in adapters (include/ewg_dune_adapters.hh)
// through Newton time steps the way NewtonMethod::apply does (newtonmethod.hh:308-517):
// linearize, error = max |residual|, proceed rule, prepare/setResidual/setMatrix/solve,
// x -= dx. Test infrastructure.
//
//   adapter_driver paired <spec.bin> <out.bin>   GpuLinearizer + GpuBiCGStabBackend
//   adapter_driver upload <spec.bin> <out.bin>   oracle linearization on the host, the
//                                                 Jacobian handed to GpuBiCGStabBackend
//   adapter_driver check                         link check only (no device needed)
//
// spec.bin (written by tests/test_adapters.py): int32 dim nx ny nz law nregions nbfaces
// has_pe nsteps; float64 lx ly lz rho[2] mu[2] g[3] dt newton_tol linear_tol
// region_params[4*nregions] K[N] phi[N] pe[N if has_pe]; int32 region[N] bc_type[B];
// float64 bc_values[2B] x0[2N]. out.bin: int32 nsteps, iterations[nsteps]; float64 x[2N].
#include "mock_dune.hh"

#include "ewg_dune_adapters.hh"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

extern "C" {
// the CPU oracle's entry points (oracle/ewo_oracle.h), used by the upload mode
struct ewo_ctx;
int ewo_create(const ewg_grid_desc *, const ewg_fluid_desc *, const ewg_material_desc *,
               const ewg_params *, ewo_ctx **);
void ewo_destroy(ewo_ctx *);
int ewo_set_static(ewo_ctx *, const double *, const double *, const int32_t *, const double *);
int ewo_set_boundary(ewo_ctx *, const int32_t *, const double *);
int ewo_set_solution(ewo_ctx *, int, const double *);
int ewo_linearize(ewo_ctx *, double, int, double *, double *);
}

using Vec = Dune::BlockVector<Dune::FieldVector<double, 2>>;

struct Problem {
    Ewg::ProblemSpec spec;
};

template <>
struct Ewg::ProblemDescription<Problem> {
    static void describe(const Problem &p, ProblemSpec &s) { s = p.spec; }
};

struct NewtonState {
    int iterations = 0;
    int numIterations() const { return iterations; }
};

struct Model {
    Vec sol[2];
    NewtonState newton;
    Vec &solution(int t) { return sol[t]; }
    const Vec &solution(int t) const { return sol[t]; }
    NewtonState &newtonMethod() { return newton; }
};

struct Simulator {
    Problem prob;
    Model mdl;
    double dt = 0.0;
    const Problem &problem() const { return prob; }
    Model &model() { return mdl; }
    double timeStepSize() const { return dt; }
};

using Linearizer =
    Ewg::GpuLinearizer<Simulator, Ewoms::MockSparseMatrixAdapter, Vec, Ewoms::NumericalIssue>;
using Backend =
    Ewg::GpuBiCGStabBackend<Simulator, Ewoms::MockSparseMatrixAdapter, Vec, Ewoms::NumericalIssue>;

template <class T>
static void rd(std::ifstream &f, T *p, size_t n) {
    f.read(reinterpret_cast<char *>(p), static_cast<std::streamsize>(n * sizeof(T)));
    if (!f) throw std::runtime_error("short spec file");
}

int main(int argc, char **argv) {
    if (argc >= 2 && std::strcmp(argv[1], "check") == 0) {
        ewg_params p;
        ewg_default_params(&p);
        std::printf("adapter_driver: linked, newton tolerance default %g\n", p.newton_tolerance);
        return 0;
    }
    if (argc < 4) {
        std::fprintf(stderr, "usage: adapter_driver paired|upload spec.bin out.bin\n");
        return 2;
    }
    const bool upload = std::strcmp(argv[1], "upload") == 0;
    Simulator sim;
    Ewg::ProblemSpec &s = sim.prob.spec;
    int32_t hdr[9];
    double dbl[14];
    std::ifstream f(argv[2], std::ios::binary);
    rd(f, hdr, 9);
    rd(f, dbl, 14);
    s.grid = {hdr[0], hdr[1], hdr[2], hdr[3], dbl[0], dbl[1], dbl[2]};
    s.fluid.rho[0] = dbl[3]; s.fluid.rho[1] = dbl[4];
    s.fluid.mu[0] = dbl[5]; s.fluid.mu[1] = dbl[6];
    for (int d = 0; d < 3; ++d) s.fluid.gravity[d] = dbl[7 + d];
    sim.dt = dbl[10];
    const double newtonTol = dbl[11], linearTol = dbl[12];
    s.law = hdr[4];
    const size_t N = (size_t)hdr[1] * hdr[2] * hdr[3], B = (size_t)hdr[6];
    const int nsteps = hdr[8];
    s.region_params.resize(4 * (size_t)hdr[5]);
    rd(f, s.region_params.data(), s.region_params.size());
    s.K.resize(N); s.phi.resize(N);
    rd(f, s.K.data(), N);
    rd(f, s.phi.data(), N);
    if (hdr[7]) { s.pe.resize(N); rd(f, s.pe.data(), N); }
    s.region.resize(N); s.bc_type.resize(B); s.bc_values.resize(2 * B);
    rd(f, s.region.data(), N);
    rd(f, s.bc_type.data(), B);
    rd(f, s.bc_values.data(), 2 * B);
    std::vector<double> x0(2 * N);
    rd(f, x0.data(), 2 * N);
    for (int t = 0; t < 2; ++t) Ewg::fromAos(x0, sim.mdl.sol[t]);

    ewg_params prm;
    ewg_default_params(&prm);
    prm.newton_tolerance = newtonTol;
    prm.linear_tolerance = linearTol;
    Linearizer lin;
    Backend backend(sim, &prm);
    ewo_ctx *ora = nullptr;
    Ewoms::MockSparseMatrixAdapter hostJ;
    Vec hostR;
    std::vector<int32_t> rowptr, colidx;
    if (upload) {
        ewg_material_desc mat{s.law, (int32_t)(s.region_params.size() / 4), s.region_params.data()};
        if (ewo_create(&s.grid, &s.fluid, &mat, &prm, &ora) != 0 ||
            ewo_set_static(ora, s.K.data(), s.phi.data(), s.region.data(),
                           s.pe.empty() ? nullptr : s.pe.data()) != 0 ||
            ewo_set_boundary(ora, s.bc_type.data(), s.bc_values.data()) != 0) {
            std::fprintf(stderr, "oracle setup failed\n");
            return 1;
        }
        Ewg::Context probe(s, prm);  // pattern of the grid
        rowptr = probe.rowptr();
        colidx = probe.colidx();
        std::vector<std::set<unsigned>> pat(N);
        for (size_t i = 0; i < N; ++i)
            for (int32_t b = rowptr[i]; b < rowptr[i + 1]; ++b) pat[i].insert((unsigned)colidx[b]);
        hostJ.reserve(pat);
        hostR.resize(N);
    } else {
        lin.init(sim, &prm);
    }

    std::vector<int32_t> its;
    std::vector<double> r(2 * N), jv(4 * (upload ? colidx.size() : 0)), xa;
    Vec dx;
    try {
        for (int step = 0; step < nsteps; ++step) {
            NewtonState &nm = sim.mdl.newton;
            nm.iterations = 0;
            double error = 1e100, lastError = 1e100;
            auto proceed = [&]() {  // newtonmethod.hh:866-884
                if (nm.iterations < 1) return true;
                if (error <= prm.newton_tolerance) return false;
                if (nm.iterations >= prm.newton_max_iterations) return error * 4.0 < lastError;
                return true;
            };
            while (proceed()) {
                lastError = error;
                const Vec *R;
                const Ewoms::MockSparseMatrixAdapter *J;
                if (upload) {
                    Ewg::toAos(sim.mdl.sol[0], xa);
                    ewo_set_solution(ora, 0, xa.data());
                    Ewg::toAos(sim.mdl.sol[1], xa);
                    ewo_set_solution(ora, 1, xa.data());
                    ewo_linearize(ora, sim.dt, nm.iterations, r.data(), jv.data());
                    Ewg::fromAos(r, hostR);
                    Ewoms::MockSparseMatrixAdapter::MatrixBlock blk;
                    for (size_t i = 0; i < N; ++i)
                        for (int32_t b = rowptr[i]; b < rowptr[i + 1]; ++b) {
                            for (int a = 0; a < 2; ++a)
                                for (int k = 0; k < 2; ++k) blk[a][k] = jv[4 * b + 2 * a + k];
                            hostJ.setBlock((unsigned)i, (unsigned)colidx[b], blk);
                        }
                    R = &hostR;
                    J = &hostJ;
                } else {
                    lin.linearize();
                    R = &lin.residual();
                    J = &lin.jacobian();
                }
                error = 0.0;
                for (size_t i = 0; i < N; ++i)
                    for (int k = 0; k < 2; ++k) error = std::max(error, std::fabs((*R)[i][k]));
                if (!proceed()) {
                    ++nm.iterations;
                    break;
                }
                backend.prepare(*J, *R);
                if (!backend.solve(dx)) throw Ewoms::NumericalIssue("linear solver did not converge");
                Vec &x = sim.mdl.sol[0];
                for (size_t i = 0; i < N; ++i)
                    for (int k = 0; k < 2; ++k) x[i][k] -= dx[i][k];
                ++nm.iterations;
            }
            if (!(error <= prm.newton_tolerance)) throw Ewoms::NumericalIssue("Newton did not converge");
            its.push_back(nm.iterations);
            sim.mdl.sol[1] = sim.mdl.sol[0];
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "adapter_driver: %s\n", e.what());
        if (ora) ewo_destroy(ora);
        return 1;
    }
    if (ora) ewo_destroy(ora);
    std::ofstream o(argv[3], std::ios::binary);
    int32_t n = (int32_t)its.size();
    o.write(reinterpret_cast<const char *>(&n), sizeof n);
    o.write(reinterpret_cast<const char *>(its.data()), (std::streamsize)(its.size() * sizeof(int32_t)));
    Ewg::toAos(sim.mdl.sol[0], xa);
    o.write(reinterpret_cast<const char *>(xa.data()), (std::streamsize)(xa.size() * sizeof(double)));
    std::printf("adapter_driver %s: %d steps, Newton iterations", argv[1], n);
    for (int v : its) std::printf(" %d", v);
    std::printf("\n");
    return 0;
}
