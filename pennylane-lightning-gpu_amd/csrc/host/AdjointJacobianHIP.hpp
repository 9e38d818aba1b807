// Adjoint Jacobian host driver and the operations container, counterparts of
// AdjointJacobianGPU (pennylane_lightning_gpu/src/algorithms/
// AdjointDiffGPU.hpp:49-597) and OpsData (the external JacobianData.hpp used
// at AdjointDiffGPU.hpp:366-374 and built by Bindings.cpp:806-840).
//
// The observable states H|lambda_j> are built here with the observables'
// applyInPlace; the reverse sweep itself runs in one C-ABI call
// (lgh_adjoint_jacobian) whose fused kernels keep the Jacobian on the device.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ObservablesHIP.hpp"
#include "StateVectorHIPManaged.hpp"

namespace Pennylane::HIP {

template <class T> class OpsData {
  public:
    using ComplexT = std::complex<T>;
    OpsData(std::vector<std::string> names, std::vector<std::vector<T>> params,
            std::vector<std::vector<std::size_t>> wires, std::vector<bool> inverses,
            std::vector<std::vector<ComplexT>> matrices)
        : names_(std::move(names)), params_(std::move(params)), wires_(std::move(wires)),
          inverses_(std::move(inverses)), matrices_(std::move(matrices)) {
        PL_HIP_ABORT_IF(names_.size() != params_.size() || names_.size() != wires_.size() ||
                            names_.size() != inverses_.size(),
                        "Inconsistent operation data");
        matrices_.resize(names_.size());
        for (const auto &p : params_) num_par_ops_ += !p.empty();
    }
    OpsData(std::vector<std::string> names, std::vector<std::vector<T>> params,
            std::vector<std::vector<std::size_t>> wires, std::vector<bool> inverses)
        : OpsData(std::move(names), std::move(params), std::move(wires), std::move(inverses),
                  std::vector<std::vector<ComplexT>>{}) {}

    std::size_t getSize() const { return names_.size(); }
    const std::vector<std::string> &getOpsName() const { return names_; }
    const std::vector<std::vector<T>> &getOpsParams() const { return params_; }
    const std::vector<std::vector<std::size_t>> &getOpsWires() const { return wires_; }
    const std::vector<bool> &getOpsInverses() const { return inverses_; }
    const std::vector<std::vector<ComplexT>> &getOpsMatrices() const { return matrices_; }
    bool hasParams(std::size_t i) const { return !params_[i].empty(); }
    std::size_t getNumParOps() const { return num_par_ops_; }

    void build(OpBuilder &b) const {
        for (std::size_t i = 0; i < names_.size(); ++i)
            b.add<T>(names_[i], wires_[i], inverses_[i], params_[i], matrices_[i]);
    }

  private:
    std::vector<std::string> names_;
    std::vector<std::vector<T>> params_;
    std::vector<std::vector<std::size_t>> wires_;
    std::vector<bool> inverses_;
    std::vector<std::vector<ComplexT>> matrices_;
    std::size_t num_par_ops_ = 0;
};

template <class T> class AdjointJacobianHIP {
  public:
    using ObsPtr = std::shared_ptr<ObservableHIP<T>>;

    OpsData<T> createOpsData(const std::vector<std::string> &names, const std::vector<std::vector<T>> &params,
                             const std::vector<std::vector<std::size_t>> &wires, const std::vector<bool> &inverses,
                             const std::vector<std::vector<std::complex<T>>> &matrices = {}) const {
        return OpsData<T>(names, params, wires, inverses, matrices);
    }

    // adjointJacobian (AdjointDiffGPU.hpp:499-596).  `sv` holds the forward
    // state (or the initial state when apply_operations).  Returns
    // jac[obs][trainable].
    std::vector<std::vector<T>> adjointJacobian(const StateVectorHIPManaged<T> &sv, const std::vector<ObsPtr> &obs,
                                                const OpsData<T> &ops, const std::vector<std::size_t> &trainable,
                                                bool apply_operations = false) const {
        PL_HIP_ABORT_IF(trainable.empty(), "No trainable parameters provided.");
        for (std::size_t i = 0; i < ops.getSize(); ++i)
            PL_HIP_ABORT_IF(ops.getOpsParams()[i].size() > 1,
                            "The operation is not supported using the adjoint differentiation method");
        OpBuilder b;
        ops.build(b);
        // lambda = U psi when requested (else psi itself, which the C ABI only
        // reads), then H_j lambda for every observable
        std::unique_ptr<StateVectorHIPManaged<T>> applied;
        if (apply_operations) {
            applied = std::make_unique<StateVectorHIPManaged<T>>(sv);
            OpBuilder fwd;
            for (std::size_t i = 0; i < ops.getSize(); ++i)
                if (!isStatePrep(ops.getOpsName()[i]))
                    fwd.add<T>(ops.getOpsName()[i], ops.getOpsWires()[i], ops.getOpsInverses()[i],
                               ops.getOpsParams()[i], ops.getOpsMatrices()[i]);
            if (fwd.size()) check(lgh_apply(applied->handle(), fwd.data(), static_cast<int64_t>(fwd.size())));
        }
        const StateVectorHIPManaged<T> &lambda = applied ? *applied : sv;
        // Observables that are words of Z factors need no state formed
        // before the sweep: the sweep forms Z_j lambda in its first pass.
        const std::size_t nq = lambda.getNumQubits();
        std::vector<uint64_t> zmasks;
        for (const auto &o : obs) {
            std::string letters;
            std::vector<std::size_t> wires;
            if (!o->pauliWord(letters, wires)) break;
            uint64_t m = 0;
            bool z_only = true;
            for (std::size_t k = 0; k < letters.size(); ++k) {
                if (letters[k] == 'Z')
                    m |= uint64_t(1) << (nq - 1 - wires[k]);
                else if (letters[k] != 'I')
                    z_only = false;
            }
            if (!z_only) break;
            zmasks.push_back(m);
        }
        const bool z_words = !obs.empty() && zmasks.size() == obs.size();
        std::vector<StateVectorHIPManaged<T>> bras;
        bras.reserve(obs.size());
        for (const auto &o : obs) {
            bras.emplace_back(lambda.getNumQubits(), lambda.getDevTag(),
                              typename StateVectorHIPManaged<T>::Uninitialized{});
            if (!z_words) o->applyFrom(lambda, bras.back());
        }
        std::vector<lgh_sv *> handles;
        for (auto &s : bras) handles.push_back(s.handle());
        std::vector<int64_t> tp(trainable.begin(), trainable.end());
        std::vector<double> jac(obs.size() * trainable.size(), 0.0);
        if (z_words)
            check(lgh_adjoint_jacobian_zwords(lambda.handle(), handles.data(), zmasks.data(),
                                              static_cast<int32_t>(handles.size()), b.data(),
                                              static_cast<int64_t>(b.size()), tp.data(),
                                              static_cast<int64_t>(tp.size()), 0, jac.data()));
        else
            check(lgh_adjoint_jacobian(lambda.handle(), handles.data(), static_cast<int32_t>(handles.size()),
                                       b.data(), static_cast<int64_t>(b.size()), tp.data(),
                                       static_cast<int64_t>(tp.size()), 0, jac.data()));
        std::vector<std::vector<T>> out(obs.size(), std::vector<T>(trainable.size()));
        for (std::size_t j = 0; j < obs.size(); ++j)
            for (std::size_t t = 0; t < trainable.size(); ++t)
                out[j][t] = static_cast<T>(jac[j * trainable.size() + t]);
        return out;
    }

    // batchAdjointJacobian (AdjointDiffGPU.hpp:392-473): the reference spreads
    // observables over GPUs with one thread each.  With one GPU the
    // observables are processed in chunks that bound the number of resident
    // observable states.
    std::vector<std::vector<T>> batchAdjointJacobian(const StateVectorHIPManaged<T> &sv,
                                                     const std::vector<ObsPtr> &obs, const OpsData<T> &ops,
                                                     const std::vector<std::size_t> &trainable,
                                                     bool apply_operations = false, std::size_t chunk = 8) const {
        std::vector<std::vector<T>> out;
        for (std::size_t b = 0; b < obs.size(); b += chunk) {
            std::vector<ObsPtr> part(obs.begin() + b, obs.begin() + std::min(obs.size(), b + chunk));
            auto j = adjointJacobian(sv, part, ops, trainable, apply_operations);
            out.insert(out.end(), j.begin(), j.end());
        }
        return out;
    }
};

} // namespace Pennylane::HIP
