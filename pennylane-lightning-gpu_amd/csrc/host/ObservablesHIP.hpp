// Observables of the MI355X statevector, the counterparts of ObservableGPU,
// NamedObsGPU, HermitianObsGPU, TensorProdObsGPU, HamiltonianGPU and
// SparseHamiltonianGPU (pennylane_lightning_gpu/src/algorithms/
// ObservablesGPU.hpp:32-610).  applyInPlace(sv) replaces sv by O sv; it is
// what the adjoint Jacobian uses to build the observable states H|lambda_j>.
#pragma once

#include <algorithm>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "StateVectorHIPManaged.hpp"

namespace Pennylane::HIP {

template <class T> std::string wiresToString(const std::vector<T> &v) {
    std::ostringstream ss;
    ss << '[';
    for (std::size_t i = 0; i < v.size(); ++i) {
        if (i) ss << ", ";
        ss << v[i];
    }
    ss << ']';
    return ss.str();
}

template <class T> class ObservableHIP {
  public:
    virtual ~ObservableHIP() = default;
    virtual void applyInPlace(StateVectorHIPManaged<T> &sv) const = 0;
    // Appends the observable as a Pauli word (letters I/X/Y/Z and their
    // wires) and returns true when it is one; then O|s> can be formed from a
    // source state in one pass (StateVectorHIPManaged::setPauliWordOf).
    virtual bool pauliWord(std::string &letters, std::vector<std::size_t> &wires) const {
        (void)letters;
        (void)wires;
        return false;
    }
    // out <- O src (out: any contents, same size as src).
    void applyFrom(const StateVectorHIPManaged<T> &src, StateVectorHIPManaged<T> &out) const {
        std::string letters;
        std::vector<std::size_t> wires;
        if (pauliWord(letters, wires)) {
            out.setPauliWordOf(src, letters, wires);
            return;
        }
        out.updateData(src, true);
        applyInPlace(out);
    }
    virtual std::string getObsName() const = 0;
    virtual std::vector<std::size_t> getWires() const = 0;
    bool operator==(const ObservableHIP &other) const {
        return typeid(*this) == typeid(other) && isEqual(other);
    }
    bool operator!=(const ObservableHIP &other) const { return !(*this == other); }

  protected:
    virtual bool isEqual(const ObservableHIP &other) const = 0;
};

// NamedObsGPU (ObservablesGPU.hpp:88-128): a named gate used as observable.
template <class T> class NamedObsHIP final : public ObservableHIP<T> {
  public:
    NamedObsHIP(std::string name, std::vector<std::size_t> wires, std::vector<T> params = {})
        : name_(std::move(name)), wires_(std::move(wires)), params_(std::move(params)) {}
    void applyInPlace(StateVectorHIPManaged<T> &sv) const override {
        sv.applyOperation(name_, wires_, false, params_);
    }
    bool pauliWord(std::string &letters, std::vector<std::size_t> &wires) const override {
        if (name_ == "Identity") return true;
        const char c = name_ == "PauliX" ? 'X' : name_ == "PauliY" ? 'Y' : name_ == "PauliZ" ? 'Z' : 0;
        if (!c || wires_.size() != 1) return false;
        letters.push_back(c);
        wires.push_back(wires_[0]);
        return true;
    }
    std::string getObsName() const override { return name_ + wiresToString(wires_); }
    std::vector<std::size_t> getWires() const override { return wires_; }
    const std::string &getName() const { return name_; }
    const std::vector<T> &getParams() const { return params_; }

  protected:
    bool isEqual(const ObservableHIP<T> &o) const override {
        const auto &x = static_cast<const NamedObsHIP &>(o);
        return name_ == x.name_ && wires_ == x.wires_ && params_ == x.params_;
    }

  private:
    std::string name_;
    std::vector<std::size_t> wires_;
    std::vector<T> params_;
};

// HermitianObsGPU (ObservablesGPU.hpp:134-180): an explicit matrix.
template <class T> class HermitianObsHIP final : public ObservableHIP<T> {
  public:
    using ComplexT = std::complex<T>;
    HermitianObsHIP(std::vector<ComplexT> matrix, std::vector<std::size_t> wires)
        : matrix_(std::move(matrix)), wires_(std::move(wires)) {
        PL_HIP_ABORT_IF(matrix_.size() != (std::size_t(1) << (2 * wires_.size())),
                        "Hermitian matrix size does not match its wires");
    }
    void applyInPlace(StateVectorHIPManaged<T> &sv) const override {
        sv.applyOperation("Hermitian", wires_, false, {}, matrix_);
    }
    std::string getObsName() const override {
        // content hash in the name, as the reference does (ObservablesGPU.hpp:168-175)
        std::size_t h = 0;
        for (const auto &z : matrix_) {
            h ^= std::hash<T>{}(z.real()) + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
            h ^= std::hash<T>{}(z.imag()) + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
        }
        return "Hermitian" + std::to_string(h);
    }
    std::vector<std::size_t> getWires() const override { return wires_; }
    const std::vector<ComplexT> &getMatrix() const { return matrix_; }

  protected:
    bool isEqual(const ObservableHIP<T> &o) const override {
        const auto &x = static_cast<const HermitianObsHIP &>(o);
        return matrix_ == x.matrix_ && wires_ == x.wires_;
    }

  private:
    std::vector<ComplexT> matrix_;
    std::vector<std::size_t> wires_;
};

// TensorProdObsGPU (ObservablesGPU.hpp:185-285): wires must be disjoint.
template <class T> class TensorProdObsHIP final : public ObservableHIP<T> {
  public:
    explicit TensorProdObsHIP(std::vector<std::shared_ptr<ObservableHIP<T>>> obs) : obs_(std::move(obs)) {
        std::set<std::size_t> seen;
        for (const auto &o : obs_)
            for (auto w : o->getWires()) {
                PL_HIP_ABORT_IF(seen.count(w), "All wires in observables must be disjoint.");
                seen.insert(w);
            }
        wires_.assign(seen.begin(), seen.end());
    }
    void applyInPlace(StateVectorHIPManaged<T> &sv) const override {
        for (const auto &o : obs_) o->applyInPlace(sv);
    }
    bool pauliWord(std::string &letters, std::vector<std::size_t> &wires) const override {
        for (const auto &o : obs_)
            if (!o->pauliWord(letters, wires)) return false;
        return true;
    }
    std::string getObsName() const override {
        std::string s;
        for (std::size_t i = 0; i < obs_.size(); ++i) {
            if (i) s += " @ ";
            s += obs_[i]->getObsName();
        }
        return s;
    }
    std::vector<std::size_t> getWires() const override { return wires_; }
    const std::vector<std::shared_ptr<ObservableHIP<T>>> &getObs() const { return obs_; }

  protected:
    bool isEqual(const ObservableHIP<T> &o) const override {
        const auto &x = static_cast<const TensorProdObsHIP &>(o);
        if (obs_.size() != x.obs_.size()) return false;
        for (std::size_t i = 0; i < obs_.size(); ++i)
            if (*obs_[i] != *x.obs_[i]) return false;
        return true;
    }

  private:
    std::vector<std::shared_ptr<ObservableHIP<T>>> obs_;
    std::vector<std::size_t> wires_;
};

// HamiltonianGPU (ObservablesGPU.hpp:290-420): sv <- sum_t c_t O_t sv.
template <class T> class HamiltonianHIP final : public ObservableHIP<T> {
  public:
    HamiltonianHIP(std::vector<T> coeffs, std::vector<std::shared_ptr<ObservableHIP<T>>> obs)
        : coeffs_(std::move(coeffs)), obs_(std::move(obs)) {
        PL_HIP_ABORT_IF(coeffs_.size() != obs_.size(), "Hamiltonian coefficients and terms differ in length");
    }
    void applyInPlace(StateVectorHIPManaged<T> &sv) const override {
        StateVectorHIPManaged<T> acc(sv.getNumQubits(), sv.getDevTag());
        acc.scale({T(0), T(0)});
        for (std::size_t t = 0; t < obs_.size(); ++t) {
            StateVectorHIPManaged<T> tmp(sv.getNumQubits(), sv.getDevTag(),
                                         typename StateVectorHIPManaged<T>::Uninitialized{});
            obs_[t]->applyFrom(sv, tmp);
            acc.axpy({coeffs_[t], T(0)}, tmp);
        }
        sv.updateData(acc);
    }
    std::string getObsName() const override {
        std::ostringstream ss;
        ss << "Hamiltonian: { 'coeffs' : " << wiresToString(coeffs_) << ", 'observables' : [";
        for (std::size_t t = 0; t < obs_.size(); ++t) {
            if (t) ss << ", ";
            ss << obs_[t]->getObsName();
        }
        ss << "]}";
        return ss.str();
    }
    std::vector<std::size_t> getWires() const override {
        std::set<std::size_t> w;
        for (const auto &o : obs_)
            for (auto x : o->getWires()) w.insert(x);
        return {w.begin(), w.end()};
    }
    const std::vector<T> &getCoeffs() const { return coeffs_; }
    const std::vector<std::shared_ptr<ObservableHIP<T>>> &getObs() const { return obs_; }

  protected:
    bool isEqual(const ObservableHIP<T> &o) const override {
        const auto &x = static_cast<const HamiltonianHIP &>(o);
        if (coeffs_ != x.coeffs_ || obs_.size() != x.obs_.size()) return false;
        for (std::size_t i = 0; i < obs_.size(); ++i)
            if (*obs_[i] != *x.obs_[i]) return false;
        return true;
    }

  private:
    std::vector<T> coeffs_;
    std::vector<std::shared_ptr<ObservableHIP<T>>> obs_;
};

// SparseHamiltonianGPU (ObservablesGPU.hpp:425-610): CSR operator kept
// resident on the device after first use (the reference re-uploads it on
// every application).
template <class T> class SparseHamiltonianHIP final : public ObservableHIP<T> {
  public:
    using ComplexT = std::complex<T>;
    using IdxT = std::conditional_t<std::is_same_v<T, float>, int32_t, int64_t>;
    SparseHamiltonianHIP(std::vector<ComplexT> data, std::vector<IdxT> indices, std::vector<IdxT> offsets,
                         std::vector<std::size_t> wires)
        : data_(std::move(data)), indices_(std::move(indices)), offsets_(std::move(offsets)),
          wires_(std::move(wires)) {
        PL_HIP_ABORT_IF(offsets_.empty(), "CSR offsets must not be empty");
        PL_HIP_ABORT_IF(data_.size() != indices_.size(), "CSR data and indices differ in length");
    }
    ~SparseHamiltonianHIP() override {
        for (auto &kv : csr_) (void)lgh_csr_destroy(kv.second);
    }
    SparseHamiltonianHIP(const SparseHamiltonianHIP &o)
        : data_(o.data_), indices_(o.indices_), offsets_(o.offsets_), wires_(o.wires_) {}

    const lgh_csr *resident(int device) const {
        auto it = csr_.find(device);
        if (it != csr_.end()) return it->second;
        std::vector<int64_t> ptr(offsets_.begin(), offsets_.end()), col(indices_.begin(), indices_.end());
        std::vector<double> val(2 * data_.size());
        for (std::size_t i = 0; i < data_.size(); ++i) {
            val[2 * i] = data_[i].real();
            val[2 * i + 1] = data_[i].imag();
        }
        lgh_csr *c = nullptr;
        check(lgh_csr_create(&c, device, StateVectorHIPManaged<T>::kDtype, static_cast<int64_t>(offsets_.size()) - 1,
                             ptr.data(), col.data(), val.data(), static_cast<int64_t>(data_.size())));
        csr_[device] = c;
        return c;
    }
    void applyInPlace(StateVectorHIPManaged<T> &sv) const override {
        PL_HIP_ABORT_IF(offsets_.size() - 1 != sv.getLength(),
                        "Sparse Hamiltonian dimension does not match the state");
        StateVectorHIPManaged<T> out(sv.getNumQubits(), sv.getDevTag());
        check(lgh_apply_csr(out.handle(), resident(sv.getDeviceID()), sv.handle()));
        sv.updateData(out);
    }
    T expval(const StateVectorHIPManaged<T> &sv) const { return sv.expvalCSR(resident(sv.getDeviceID())); }
    std::string getObsName() const override {
        std::ostringstream ss;
        ss << "SparseHamiltonian: {\n'data' : ";
        for (const auto &d : data_) ss << d;
        ss << ",\n'indices' : ";
        for (const auto &i : indices_) ss << i;
        ss << ",\n'offsets' : ";
        for (const auto &o : offsets_) ss << o;
        ss << "\n}";
        return ss.str();
    }
    std::vector<std::size_t> getWires() const override { return wires_; }

  protected:
    bool isEqual(const ObservableHIP<T> &o) const override {
        const auto &x = static_cast<const SparseHamiltonianHIP &>(o);
        return data_ == x.data_ && indices_ == x.indices_ && offsets_ == x.offsets_ && wires_ == x.wires_;
    }

  private:
    std::vector<ComplexT> data_;
    std::vector<IdxT> indices_;
    std::vector<IdxT> offsets_;
    std::vector<std::size_t> wires_;
    mutable std::map<int, lgh_csr *> csr_;
};

} // namespace Pennylane::HIP
