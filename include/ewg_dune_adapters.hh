// ewg_dune_adapters.hh -- C++ drop-in adapters over the ewg C ABI (include/ewg_capi.h).
//
// They give the GPU library the method sets of the two eWoms plug-in points it replaces,
// so a TypeTag can select them through its properties:
//
//   GpuLinearizer      for property Linearizer
//                      (FvBaseLinearizer, ewoms/numerics/discretizations/common/
//                       fvbaselinearizer.hh:115-256; selected at fvbasediscretization.hh:204)
//   GpuBiCGStabBackend for the LinearSolverSplice / LinearSolverBackend
//                      (ParallelBiCGStabSolverBackend, ewoms/numerics/linear/
//                       parallelbicgstabbackend.hh:81-163, parallelbasebackend.hh:202-344)
//
// Header only, with no Dune include: the simulator, the sparse-matrix adapter, the
// vector type and the exception type for numerical issues are template parameters, and
// the types only need the members the reference's own callers use:
//   Simulator:     problem(), model().solution(timeIdx) (indexable, x[i][pvIdx]),
//                  timeStepSize(), model().newtonMethod().numIterations()
//   SparseMatrix:  reserve(std::vector<std::set<unsigned>>), setBlock(row, col, block),
//                  block(row, col), typedef MatrixBlock (2x2, block[i][j])
//                  (IstlSparseMatrixAdapter, linear/istlsparsematrixadapter.hh:72-164)
//   Vector:        resize(n), size(), v[i][k] (Dune::BlockVector<FieldVector<double,2>>)
// A problem plugs in by specialising Ewg::ProblemDescription<Problem> (the GPU needs the
// grid, fluid, material and boundary data as arrays, where the reference asks the
// problem per element).
//
// When the backend is given the matrix of a GpuLinearizer, the system never leaves the
// device (the linearizer stops copying the Jacobian to the host from then on); given any
// other matrix, the backend uploads it, like ParallelBaseBackend::setMatrix copies it.
#ifndef EWG_DUNE_ADAPTERS_HH
#define EWG_DUNE_ADAPTERS_HH

#include "ewg_capi.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace Ewg {

// Static description of a problem on the structured grid (see ewg_create, ewg_set_static,
// ewg_set_boundary for the meaning and order of every field).
struct ProblemSpec {
    ewg_grid_desc grid{};
    ewg_fluid_desc fluid{};
    int32_t law = EWG_LAW_VAN_GENUCHTEN;
    std::vector<double> region_params;  // 4 per region
    std::vector<double> K, phi, pe;     // per cell; pe empty unless Brooks-Corey per cell
    std::vector<int32_t> region;        // per cell
    std::vector<int32_t> bc_type;       // per boundary face, sides x-, x+, y-, y+, z-, z+
    std::vector<double> bc_values;      // 2 per boundary face
};

// Specialise for a problem type: static void describe(const Problem&, ProblemSpec&).
template <class Problem>
struct ProblemDescription;

// Status > 0 of the C ABI is a numerical issue (NumericalIssue is thrown), status < 0 a
// misuse or device error (std::runtime_error).
template <class NumericalIssue>
inline void check(ewg_ctx *ctx, int rc, const char *what) {
    if (rc == EWG_OK) return;
    std::string msg = std::string(what) + ": " + (ctx ? ewg_last_error(ctx) : "no context");
    if (rc > 0) throw NumericalIssue(msg);
    throw std::runtime_error(msg);
}

// One device context built from a ProblemSpec.
class Context {
public:
    Context(const ProblemSpec &spec, const ewg_params &params) {
        ewg_material_desc mat{};
        mat.law = spec.law;
        mat.nregions = static_cast<int32_t>(spec.region_params.size() / 4);
        mat.region_params = spec.region_params.data();
        int rc = ewg_create(&spec.grid, &spec.fluid, &mat, &params, &ctx_);
        if (rc != EWG_OK || !ctx_) throw std::runtime_error("ewg_create failed");
        N_ = static_cast<size_t>(ewg_num_cells(ctx_));
        std::vector<int32_t> region = spec.region;
        if (region.empty()) region.assign(N_, 0);
        rc = ewg_set_static(ctx_, spec.K.data(), spec.phi.data(), region.data(),
                            spec.pe.empty() ? nullptr : spec.pe.data());
        if (rc == EWG_OK) rc = ewg_set_boundary(ctx_, spec.bc_type.data(), spec.bc_values.data());
        if (rc != EWG_OK) {
            std::string msg = ewg_last_error(ctx_);
            ewg_destroy(ctx_);
            throw std::runtime_error("ewg static data: " + msg);
        }
        rowptr_.resize(N_ + 1);
        colidx_.resize(static_cast<size_t>(ewg_num_blocks(ctx_)));
        ewg_get_pattern(ctx_, rowptr_.data(), colidx_.data());
    }
    ~Context() {
        if (ctx_) ewg_destroy(ctx_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    ewg_ctx *get() const { return ctx_; }
    size_t numCells() const { return N_; }
    const std::vector<int32_t> &rowptr() const { return rowptr_; }
    const std::vector<int32_t> &colidx() const { return colidx_; }

private:
    ewg_ctx *ctx_ = nullptr;
    size_t N_ = 0;
    std::vector<int32_t> rowptr_, colidx_;
};

// Matrices whose values live on a GpuLinearizer's device context, by address.
struct DeviceSystems {
    struct Entry {
        ewg_ctx *ctx;
        bool *hostJacobianWanted;
    };
    static DeviceSystems &instance() {
        static DeviceSystems s;
        return s;
    }
    void add(const void *matrix, Entry e) {
        std::lock_guard<std::mutex> g(m_);
        map_[matrix] = e;
    }
    void remove(const void *matrix) {
        std::lock_guard<std::mutex> g(m_);
        map_.erase(matrix);
    }
    bool find(const void *matrix, Entry &e) {
        std::lock_guard<std::mutex> g(m_);
        auto it = map_.find(matrix);
        if (it == map_.end()) return false;
        e = it->second;
        return true;
    }

private:
    std::mutex m_;
    std::map<const void *, Entry> map_;
};

template <class Vector>
inline void toAos(const Vector &v, std::vector<double> &out) {
    out.resize(2 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        out[2 * i] = v[i][0];
        out[2 * i + 1] = v[i][1];
    }
}
template <class Vector>
inline void fromAos(const std::vector<double> &in, Vector &v) {
    if (v.size() != in.size() / 2) v.resize(in.size() / 2);
    for (size_t i = 0; i < v.size(); ++i) {
        v[i][0] = in[2 * i];
        v[i][1] = in[2 * i + 1];
    }
}

// FvBaseLinearizer's method set (fvbaselinearizer.hh:115-256) on the GPU.
template <class Simulator, class SparseMatrixAdapter, class GlobalEqVector, class NumericalIssue>
class GpuLinearizer {
public:
    GpuLinearizer() = default;
    ~GpuLinearizer() { DeviceSystems::instance().remove(&jacobian_); }
    GpuLinearizer(const GpuLinearizer &) = delete;
    GpuLinearizer &operator=(const GpuLinearizer &) = delete;

    static void registerParameters() {}

    // FvBaseLinearizer::init (fvbaselinearizer.hh:127-131)
    void init(Simulator &simulator, const ewg_params *params = nullptr) {
        simulator_ = &simulator;
        ProblemSpec spec;
        ProblemDescription<typename std::decay<decltype(simulator.problem())>::type>::describe(
            simulator.problem(), spec);
        ewg_params p;
        ewg_default_params(&p);
        if (params) p = *params;
        ctx_.reset(new Context(spec, p));
        // createMatrix_ (fvbaselinearizer.hh:303-340): ascending columns, diagonal included
        const size_t N = ctx_->numCells();
        std::vector<std::set<unsigned>> pattern(N);
        for (size_t i = 0; i < N; ++i)
            for (int32_t b = ctx_->rowptr()[i]; b < ctx_->rowptr()[i + 1]; ++b)
                pattern[i].insert(static_cast<unsigned>(ctx_->colidx()[b]));
        jacobian_.reserve(pattern);
        residual_.resize(N);
        DeviceSystems::instance().add(&jacobian_, {ctx_->get(), &hostJacobianWanted_});
    }

    void eraseMatrix() {}

    // FvBaseLinearizer::linearize (fvbaselinearizer.hh:150-160)
    void linearize() {
        linearizeDomain();
        linearizeAuxiliaryEquations();
        finalize();
    }

    // FvBaseLinearizer::linearizeDomain (fvbaselinearizer.hh:166-197): the current and
    // previous solutions go to the device, the residual (and, unless a GPU backend takes
    // the matrix from the device, the Jacobian) come back in the reference's layouts
    void linearizeDomain() {
        ewg_ctx *c = ctx_->get();
        auto &model = simulator_->model();
        toAos(model.solution(/*timeIdx=*/0), x_);
        check<NumericalIssue>(c, ewg_set_solution(c, 0, x_.data()), "ewg_set_solution");
        toAos(model.solution(/*timeIdx=*/1), x_);
        check<NumericalIssue>(c, ewg_set_solution(c, 1, x_.data()), "ewg_set_solution");
        const int newtonIter = static_cast<int>(model.newtonMethod().numIterations());
        resid_.resize(2 * ctx_->numCells());
        jac_.resize(hostJacobianWanted_ ? 4 * ctx_->colidx().size() : 0);
        check<NumericalIssue>(c, ewg_linearize(c, simulator_->timeStepSize(), newtonIter,
                                               resid_.data(),
                                               hostJacobianWanted_ ? jac_.data() : nullptr),
                              "ewg_linearize");
        fromAos(resid_, residual_);
        if (hostJacobianWanted_) {
            typename SparseMatrixAdapter::MatrixBlock blk;
            const size_t N = ctx_->numCells();
            for (size_t i = 0; i < N; ++i)
                for (int32_t b = ctx_->rowptr()[i]; b < ctx_->rowptr()[i + 1]; ++b) {
                    for (int r = 0; r < 2; ++r)
                        for (int k = 0; k < 2; ++k) blk[r][k] = jac_[4 * b + 2 * r + k];
                    jacobian_.setBlock(static_cast<unsigned>(i),
                                       static_cast<unsigned>(ctx_->colidx()[b]), blk);
                }
        }
    }

    void linearizeAuxiliaryEquations() {}
    void finalize() {}

    SparseMatrixAdapter &jacobian() { return jacobian_; }
    const SparseMatrixAdapter &jacobian() const { return jacobian_; }
    GlobalEqVector &residual() { return residual_; }
    const GlobalEqVector &residual() const { return residual_; }

    // no constraint DOFs on this path (fvbaselinearizer.hh:256)
    const std::map<unsigned, int> &constraintsMap() const { return constraints_; }

    ewg_ctx *deviceContext() const { return ctx_ ? ctx_->get() : nullptr; }

private:
    Simulator *simulator_ = nullptr;
    std::unique_ptr<Context> ctx_;
    SparseMatrixAdapter jacobian_;
    GlobalEqVector residual_;
    std::map<unsigned, int> constraints_;
    std::vector<double> x_, resid_, jac_;
    bool hostJacobianWanted_ = true;
};

// ParallelBiCGStabSolverBackend's method set (parallelbasebackend.hh:202-344) on the GPU.
template <class Simulator, class SparseMatrixAdapter, class Vector, class NumericalIssue>
class GpuBiCGStabBackend {
public:
    explicit GpuBiCGStabBackend(const Simulator &simulator, const ewg_params *params = nullptr)
        : simulator_(&simulator) {
        ewg_default_params(&params_);
        if (params) params_ = *params;
    }

    static void registerParameters() {}

    void eraseMatrix() {}

    // prepare (parallelbasebackend.hh:231-270): a GpuLinearizer's matrix is used where it
    // is; for any other matrix the backend keeps its own context and uploads
    void prepare(const SparseMatrixAdapter &M, const Vector &b) {
        DeviceSystems::Entry e{};
        if (DeviceSystems::instance().find(&M, e)) {
            ctx_ = e.ctx;
            *e.hostJacobianWanted = false;
            upload_ = false;
        } else {
            if (!own_) {
                ProblemSpec spec;
                ProblemDescription<typename std::decay<decltype(simulator_->problem())>::type>::
                    describe(simulator_->problem(), spec);
                own_.reset(new Context(spec, params_));
            }
            ctx_ = own_->get();
            upload_ = true;
        }
        ewg_set_params(ctx_, &params_);
        setResidual(b);
        setMatrix(M);
    }

    // setResidual / getResidual (parallelbasebackend.hh:272-289)
    void setResidual(const Vector &b) {
        toAos(b, b_);
        if (upload_) check<NumericalIssue>(ctx_, ewg_set_system(ctx_, b_.data(), nullptr), "ewg_set_system");
    }
    void getResidual(Vector &b) const { fromAos(b_, b); }

    // setMatrix (parallelbasebackend.hh:291-312): the values in BCRS order
    void setMatrix(const SparseMatrixAdapter &M) {
        if (!upload_) return;
        const auto &rp = own_->rowptr();
        const auto &ci = own_->colidx();
        vals_.resize(4 * ci.size());
        for (size_t i = 0; i + 1 < rp.size(); ++i)
            for (int32_t b = rp[i]; b < rp[i + 1]; ++b) {
                const auto &blk = M.block(static_cast<unsigned>(i), static_cast<unsigned>(ci[b]));
                for (int r = 0; r < 2; ++r)
                    for (int k = 0; k < 2; ++k) vals_[4 * b + 2 * r + k] = blk[r][k];
            }
        check<NumericalIssue>(ctx_, ewg_set_system(ctx_, nullptr, vals_.data()), "ewg_set_system");
    }

    // solve (parallelbasebackend.hh:314-338): false when not converged, NumericalIssue on
    // breakdown or a failed preconditioner
    bool solve(Vector &x) {
        int32_t its = 0;
        x_.resize(b_.size());
        int rc = ewg_solve(ctx_, x_.data(), &its);
        iterations_ = static_cast<size_t>(its);
        if (rc == EWG_NOT_CONVERGED) return false;
        check<NumericalIssue>(ctx_, rc, "ewg_solve");
        fromAos(x_, x);
        return true;
    }

    size_t iterations() const { return iterations_; }

private:
    const Simulator *simulator_;
    ewg_params params_{};
    std::unique_ptr<Context> own_;
    ewg_ctx *ctx_ = nullptr;
    bool upload_ = true;
    std::vector<double> b_, x_, vals_;
    size_t iterations_ = 0;
};

}  // namespace Ewg

#endif  // EWG_DUNE_ADAPTERS_HH
